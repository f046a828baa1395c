"""Loader for the native library vine_amd/libvine.so (HIP kernels + C ABI).

There is no fallback: if the library has not been built the import fails with
an error that says how to build it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvine.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "vine_amd native library missing at %s; build it with "
                "`python -m vine_amd.build`" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
c_vp = ctypes.c_void_p
P = ctypes.POINTER


def _sig(L, name, res, args):
    fn = getattr(L, name)
    fn.restype = res
    fn.argtypes = args


def _declare(L):
    pp = P(c_vp)
    _sig(L, "vine_gen_grid", c_i32, [c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_u64, pp])
    _sig(L, "vine_gen_knn", c_i32, [c_i64, c_i32, c_f64, c_u64, pp])
    _sig(L, "vine_gen_rmat", c_i32, [c_i32, c_i32, c_f64, c_f64, c_f64, c_i32, c_i32, c_u64, pp])
    _sig(L, "vine_host_graph_from_csr", c_i32, [c_i32, c_i64, c_vp, c_vp, c_vp, c_vp, pp])
    _sig(L, "vine_host_graph_free", None, [c_vp])
    _sig(L, "vine_host_graph_n", c_i32, [c_vp])
    _sig(L, "vine_host_graph_m", c_i64, [c_vp])
    for nm in ("vine_host_graph_row_ptr", "vine_host_graph_dst",
               "vine_host_graph_len", "vine_host_graph_delay"):
        _sig(L, nm, c_vp, [c_vp])
    _sig(L, "vine_min_delay", c_i64, [c_vp, c_i32, c_i32])
    _sig(L, "vine_gen_queries_random", c_i32,
         [c_vp, c_i32, c_i32, c_i32, c_i32, c_u64, c_i32, c_vp, c_vp, c_vp, c_vp])
    _sig(L, "vine_gen_queries_hops", c_i32,
         [c_vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_u64, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp])
    _sig(L, "vine_gen_preset_graph", c_i32, [c_i32, c_u64, pp])
    _sig(L, "vine_read_graph", c_i32, [ctypes.c_char_p, c_i32, pp])
    _sig(L, "vine_read_queries", c_i64, [ctypes.c_char_p, c_i32, c_i64, c_vp, c_vp, c_vp, c_vp])
    _sig(L, "vine_last_error", ctypes.c_char_p, [])
    if hasattr(L, "vine_declare_solver"):
        pass
