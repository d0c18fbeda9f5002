"""Drop-in for the reference's chain_caller.py (chain_caller.py:23-52).

The Chain functions keep the reference's names and argument order and run on
the MI355X engine:

* chain_gpu(coords (D, N), eb, p, threshold, diff) -> bytes (LCH1)
  replaces lcps_gpu_cpp.lcps_compress2_gpu (cu/chain/lcps_gpu_cpp.cu:1095)
* gpu_decompress(bytes) -> float64 (N, D); float32 for LCG2
  replaces lcps_gpu_decompress_ans.lcps_decompress_gpu
  (cu/decompress/lcps_gpu_decompress_ans.cu:2898): LCG2 streams of the
  reference GPU compressor decode bit for bit like its lcg2::decode when their
  components use the Fixed encoding (Zstd sections, inflated on the host);
  Tagged and Raw components are coded with nvCOMP ANS and are refused
* chain(coords (N, D), eb, p, threshold, diff, bl, l1, l2, ec) -> [bytes, stats...]
  mirrors lcps_cpp.lcps_compress2 (cpp/chain/lcps_cpp.cpp:2340): ec=False
  returns the reference's LCP3 bytes exactly; ec=True returns the reference's
  own LCP2 container, so streams stored by a drop-in caller stay readable by
  the reference's lcps_cpu_decompress
* chain_decompress(bytes) -> float64 (N, D) for LCH1, LCP3 and LCP2 streams
  (LCP2 is converted to LCP3 on the host first)

With ec=True the GPU's LCP3 streams go through the host-side Zstd-22 / RVN1
chooser (chaincompressor_amd.lcp2), byte-identical to
lcps_compress2(..., ec=True) with the reference's libzstd. Zstd-22 is a serial
host codec; chain(..., container="lch1") returns this engine's entropy-coded
LCH1 instead (entropy coding on the GPU, decoded by chain_decompress).

The LCP-S and Draco wrappers of the reference (lcps_compress,
lcps_decompress, dracoIds) are competitor baselines outside this engine.
"""
from chaincompressor_amd import lcps_gpu_cpp, lcps_gpu_decompress_ans


def chain(coords, eb, p, threshold, diff, bl=True, l1=True, l2=True, ec=True, *, container="lcp2"):
    if container not in ("lch1", "lcp2"):
        raise ValueError("container must be 'lch1' or 'lcp2'")
    if ec and container == "lcp2":
        from chaincompressor_amd import lcp2
        raw = lcps_gpu_cpp.lcps_compress2(coords, eb, p, threshold, diff, bl, l1, l2, False)
        out, stats = lcp2.to_lcp2(raw[0])
        return [out] + stats
    return lcps_gpu_cpp.lcps_compress2(coords, eb, p, threshold, diff, bl, l1, l2, ec)


def _device_stream(data):
    """LCP2 (the reference CPU compressor's container) becomes the LCP3 of the
    same streams on the host, LCG2 (its GPU compressor's) gets its Zstd
    sections inflated; LCH1 and LCP3 go to the GPU as they are."""
    b = bytes(data)
    if b[:4] == b"LCP2":
        from chaincompressor_amd import lcp2
        return lcp2.to_lcp3(b)
    if b[:4] == b"LCG2":
        from chaincompressor_amd import lcp2
        return lcp2.inflate_lcg2(b)
    return b


def chain_decompress(compressed_data, timer=False):
    return lcps_gpu_decompress_ans.lcps_decompress_gpu(_device_stream(compressed_data))


def gpu_decompress(compressed_data, timer=False):
    return lcps_gpu_decompress_ans.lcps_decompress_gpu(_device_stream(compressed_data))


def chain_gpu(coords, eb, p, threshold, diff, timer=False):
    return lcps_gpu_cpp.lcps_compress2_gpu(coords, eb, p, threshold, diff)


def lcps_compress(coords, eb, reorder=False):
    raise NotImplementedError("LCP-S is a competitor baseline of the reference, not part of the Chain engine")


def lcps_decompress(compressed):
    raise NotImplementedError("LCP-S is a competitor baseline of the reference, not part of the Chain engine")


def dracoIds(data, qb):
    raise NotImplementedError("Draco is a competitor baseline of the reference, not part of the Chain engine")
