// rANS helpers of the Chain engine's kernels.
//
// The engine's RVC1 sections use the normalised value-rANS of the reference
// (RVN1, cpp/chain/lcps_cpp.cpp:539-821): same precision rule, same table
// normalisation, state interval [4M, 1024M) and byte renormalisation. The
// reference normalises in x87 80-bit long double; chain_int.hpp reproduces
// those roundings exactly in integer arithmetic so the table can be built on
// the GPU (and by host tools).
#pragma once
#include <cstdint>

#include "chain_common.hpp"
#include "chain_int.hpp"
