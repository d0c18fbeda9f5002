// pybind11 modules over the C ABI of include/chain_engine.h.
//
// Built twice from this file:
//   -DCHAIN_MODULE_COMPRESS   -> lcps_gpu_cpp            (cu/chain/lcps_gpu_cpp.cu:1095-1102)
//   -DCHAIN_MODULE_DECOMPRESS -> lcps_gpu_decompress_ans (cu/decompress/lcps_gpu_decompress_ans.cu:2898-2921)
// The names, argument order and defaults follow the reference modules; the
// GIL is released around every engine call. Differences from the reference
// GPU modules, kept on purpose for bit-exactness with the CPU path:
//   * eb is taken as a Python float (double), not float32;
//   * quantisation is FP64 with exact corrections (LCP2 semantics);
//   * decompression returns float64 unless dtype="float32" is asked for;
//     LCG2 streams (the reference GPU compressor's) return float32, as the
//     reference does, with their Zstd sections inflated on the host
//     (libchain_lcp2, built with CHAIN_WITH_LCG2) and Fixed components only.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/chain_engine.h"
#ifdef CHAIN_WITH_LCG2
#include "../../include/chain_lcp2.h"
#endif

namespace py = pybind11;

namespace {

chain_engine* g_engine = nullptr;
std::mutex g_mu;

chain_engine* engine() {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_engine) {
        const char* dev = std::getenv("CHAIN_DEVICE");
        if (chain_engine_create(dev ? std::atoi(dev) : 0, 0, &g_engine) != CHAIN_OK)
            throw std::runtime_error(std::string("chain engine: ") + chain_last_error(nullptr));
    }
    return g_engine;
}

void check(chain_engine* e, int rc) {
    if (rc == CHAIN_OK) return;
    const std::string msg = chain_last_error(e);
    if (rc == CHAIN_EINVAL) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

struct Points {
    py::array arr;
    int dtype;
};

Points as_points(py::object obj) {
    py::array a = py::array::ensure(obj);
    if (!a) throw py::value_error("points cannot be converted to an ndarray");
    if (a.dtype().is(py::dtype::of<float>())) {
        return {py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(a), CHAIN_F32};
    }
    return {py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(a), CHAIN_F64};
}

py::bytes take(chain_buf& b) {
    py::bytes out(reinterpret_cast<const char*>(b.data), b.size);
    chain_buf_free(&b);
    return out;
}

}  // namespace

#ifdef CHAIN_MODULE_COMPRESS

// lcps_compress2_gpu(points (D, N), eb, p, threshold, diff) -> bytes (LCH1)
static py::bytes lcps_compress2_gpu(py::object points, double eb, int64_t p, int64_t threshold, int64_t diff) {
    (void)diff;   // accepted and ignored, as in cu/chain/lcps_gpu_cpp.cu:721
    Points pts = as_points(points);
    if (pts.arr.ndim() != 2) throw py::value_error("points must be a 2-D ndarray with shape (D, N)");
    const uint32_t dims = (uint32_t)pts.arr.shape(0);
    const uint64_t n = (uint64_t)pts.arr.shape(1);
    chain_engine* e = engine();
    chain_buf out{nullptr, 0};
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = chain_compress(e, pts.arr.data(), pts.dtype, CHAIN_SOA, dims, n, eb, p, threshold, 0, CHAIN_FMT_LCH1,
                            &out, nullptr);
    }
    check(e, rc);
    return take(out);
}

// lcps_compress2(points (N, D), eb, p, threshold, diff, bl, l1, l2, ec) -> [bytes, stats...]
// (cpp/chain/lcps_cpp.cpp:2340-2355): ec=False writes the reference's LCP3
// byte for byte; ec=True writes this engine's entropy-coded LCH1.
static py::list lcps_compress2(py::object points, double eb, int64_t p, int64_t threshold, int64_t diff, bool bl,
                               bool l1, bool l2, bool ec) {
    (void)diff;
    Points pts = as_points(points);
    if (pts.arr.ndim() != 2) throw py::value_error("points must be a 2-D ndarray with shape (N, D)");
    const uint64_t n = (uint64_t)pts.arr.shape(0);
    const uint32_t dims = (uint32_t)pts.arr.shape(1);
    const uint32_t flags = (bl ? 0u : CHAIN_NO_BL) | (l1 ? 0u : CHAIN_NO_L1) | (l2 ? 0u : CHAIN_NO_L2);
    chain_engine* e = engine();
    chain_buf out{nullptr, 0};
    chain_stats st{};
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = chain_compress(e, pts.arr.data(), pts.dtype, CHAIN_AOS, dims, n, eb, p, threshold, flags,
                            ec ? CHAIN_FMT_LCH1 : CHAIN_FMT_LCP3, &out, &st);
    }
    check(e, rc);
    py::list ret;
    ret.append(take(out));
    ret.append(st.block_replacement_size);
    ret.append(st.idrel_size);
    ret.append(st.rel_size);
    ret.append(st.num_chained);
    return ret;
}

PYBIND11_MODULE(lcps_gpu_cpp, m) {
    m.doc() = "MI355X Chain compressor (HIP, gfx950): LCH1 containers with chunked interleaved rANS.";
    m.def("lcps_compress2_gpu", &lcps_compress2_gpu, py::arg("points"), py::arg("eb"), py::arg("p"),
          py::arg("threshold"), py::arg("diff"));
    m.def("lcps_compress2", &lcps_compress2, py::arg("points"), py::arg("eb"), py::arg("p"), py::arg("threshold"),
          py::arg("diff"), py::arg("bl") = true, py::arg("l1") = true, py::arg("l2") = true, py::arg("ec") = true);
}

#endif

#ifdef CHAIN_MODULE_DECOMPRESS

// lcps_decompress_gpu(compressed_data, output_soa=False, dtype=None)
static py::array lcps_decompress_gpu(py::bytes compressed, bool output_soa, py::object dtype_arg) {
    std::string s = compressed;
    const bool lcg2 = s.size() >= 4 && std::memcmp(s.data(), "LCG2", 4) == 0;
    const std::string dtype = dtype_arg.is_none() ? (lcg2 ? "float32" : "float64") : py::str(dtype_arg).cast<std::string>();
    if (dtype != "float64" && dtype != "float32") throw py::value_error("dtype must be 'float64' or 'float32'");
    if (lcg2) {
#ifdef CHAIN_WITH_LCG2
        uint8_t* inflated = nullptr;
        size_t len = 0;
        char err[256] = {0};
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = chain_lcg2_inflate(reinterpret_cast<const uint8_t*>(s.data()), s.size(), &inflated, &len, err,
                                    sizeof err);
        }
        if (rc != CHAIN_LCP2_OK) throw std::runtime_error(std::string("LCG2: ") + err);
        s.assign(reinterpret_cast<const char*>(inflated), len);
        chain_lcp2_free(inflated);
#else
        throw std::runtime_error("LCG2 streams need libchain_lcp2 (libzstd) to inflate their Zstd sections");
#endif
    }
    const int odt = dtype == "float32" ? CHAIN_F32 : CHAIN_F64;
    chain_engine* e = engine();
    chain_buf out{nullptr, 0};
    uint64_t n = 0;
    uint32_t dims = 0;
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = chain_decompress(e, reinterpret_cast<const uint8_t*>(s.data()), s.size(), odt,
                              output_soa ? CHAIN_SOA : CHAIN_AOS, &out, &n, &dims);
    }
    check(e, rc);
    std::vector<py::ssize_t> shape = output_soa ? std::vector<py::ssize_t>{(py::ssize_t)dims, (py::ssize_t)n}
                                                : std::vector<py::ssize_t>{(py::ssize_t)n, (py::ssize_t)dims};
    // the array takes ownership of the library's buffer (no copy), as the
    // reference decoder hands its storage over with a capsule
    uint8_t* data = out.data;
    py::capsule owner(data, [](void* ptr) {
        chain_buf b{static_cast<uint8_t*>(ptr), 0};
        chain_buf_free(&b);
    });
    return odt == CHAIN_F32 ? py::array(py::dtype::of<float>(), shape, data, owner)
                            : py::array(py::dtype::of<double>(), shape, data, owner);
}

PYBIND11_MODULE(lcps_gpu_decompress_ans, m) {
    m.doc() = "MI355X Chain decompressor (HIP, gfx950) for LCH1, LCP3 and LCG2 (Fixed components) streams.";
    m.def("lcps_decompress_gpu", &lcps_decompress_gpu, py::arg("compressed_data"), py::arg("output_soa") = false,
          py::arg("dtype") = py::none());
}

#endif
