"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this.
The oracle restates the reference kernels single-threaded (fz.cu:49-209,
lorenzo_var.cuh:47-597); ``nthreads > 1`` runs the same code under OpenMP.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MODE_FIXED = 0
MODE_LITERAL = 1
MODE_OUTLIER = 2


class Info(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("ndim", ctypes.c_int32),
        ("vmin", ctypes.c_float), ("vmax", ctypes.c_float), ("range", ctypes.c_float),
        ("ebx2_r", ctypes.c_float), ("ebx2", ctypes.c_float),
        ("eb_abs", ctypes.c_double),
        ("len", ctypes.c_uint64), ("T", ctypes.c_uint64), ("W", ctypes.c_uint64), ("N", ctypes.c_uint64),
        ("K", ctypes.c_uint64), ("payload_bytes", ctypes.c_uint64), ("tail_len", ctypes.c_uint64),
        ("overflow", ctypes.c_uint64), ("n_raw", ctypes.c_uint64), ("side_bytes", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _Res(ctypes.Structure):
    _fields_ = [("min", ctypes.c_double), ("max", ctypes.c_double), ("rng", ctypes.c_double),
                ("std", ctypes.c_double)]


class _Reduced(ctypes.Structure):
    _fields_ = [("PSNR", ctypes.c_double), ("MSE", ctypes.c_double), ("NRMSE", ctypes.c_double),
                ("coeff", ctypes.c_double)]


class _MaxErr(ctypes.Structure):
    _fields_ = [("abs", ctypes.c_double), ("rel", ctypes.c_double), ("pwrrel", ctypes.c_double),
                ("idx", ctypes.c_uint64)]


class _Autocor(ctypes.Structure):
    _fields_ = [("lag_one", ctypes.c_double), ("lag_two", ctypes.c_double)]


class Stats(ctypes.Structure):
    """cusz_stats (type.h:185-194) with size_t as uint64; same layout as fz_stats."""
    _fields_ = [("odata", _Res), ("xdata", _Res), ("reduced", _Reduced), ("max_err", _MaxErr),
                ("autocor", _Autocor), ("user_eb", ctypes.c_double), ("len", ctypes.c_uint64)]

    def as_dict(self):
        def sub(st):
            return {k: getattr(st, k) for k, _ in st._fields_}
        return {"odata": sub(self.odata), "xdata": sub(self.xdata), "reduced": sub(self.reduced),
                "max_err": sub(self.max_err), "len": self.len}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        u32, u64, f32, f64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_int
        vp = ctypes.c_void_p
        L.fzo_ndim.argtypes = [u32, u32, u32]
        L.fzo_ndim.restype = i32
        for f in ("fzo_tiles", "fzo_flag_words", "fzo_payload_max_bytes"):
            getattr(L, f).argtypes = [u64]
            getattr(L, f).restype = u64
        L.fzo_range.argtypes = [vp, u64, ctypes.POINTER(f32), ctypes.POINTER(f32)]
        L.fzo_range.restype = i32
        L.fzo_eb_params.argtypes = [f32, f64, ctypes.POINTER(f64), ctypes.POINTER(f32), ctypes.POINTER(f32)]
        L.fzo_eb_params.restype = None
        L.fzo_dualquant.argtypes = [vp, vp, u32, u32, u32, f32, i32, i32]
        L.fzo_dualquant.restype = u64
        L.fzo_bitshuffle_flags.argtypes = [vp, vp, u64, vp, vp, i32]
        L.fzo_bitshuffle_flags.restype = None
        L.fzo_exclusive_sum.argtypes = [vp, vp, u64, i32]
        L.fzo_exclusive_sum.restype = None
        L.fzo_encode.argtypes = [vp, vp, vp, u64, i32]
        L.fzo_encode.restype = None
        L.fzo_unshuffle.argtypes = [vp, vp, u64, i32]
        L.fzo_unshuffle.restype = None
        L.fzo_reconstruct.argtypes = [vp, vp, u32, u32, u32, f32, i32]
        L.fzo_reconstruct.restype = None
        L.fzo_compress.argtypes = [vp, u64, u32, u32, u32, f64, i32, i32, vp, vp, ctypes.POINTER(Info),
                                   vp, vp, vp, vp]
        L.fzo_compress.restype = i32
        L.fzo_decompress.argtypes = [vp, u64, vp, u64, u64, u32, u32, u32, f32, i32, vp, vp]
        L.fzo_decompress.restype = i32
        L.fzo_outlier_max_bytes.argtypes = [u64]
        L.fzo_outlier_max_bytes.restype = u64
        L.fzo_compress_outlier.argtypes = [vp, u64, u32, u32, u32, f64, vp, vp, vp, u64, ctypes.POINTER(Info), vp]
        L.fzo_compress_outlier.restype = i32
        L.fzo_decompress_outlier.argtypes = [vp, u64, vp, u64, vp, u64, u64, u32, u32, u32, f32, f32, vp]
        L.fzo_decompress_outlier.restype = i32
        L.fzo_verify.argtypes = [vp, vp, u64, ctypes.POINTER(Stats)]
        L.fzo_verify.restype = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def ndim(x, y, z):
    return lib().fzo_ndim(x, y, z)


def tiles(n):
    return lib().fzo_tiles(n)


def flag_words(n):
    return lib().fzo_flag_words(n)


def payload_max_bytes(n):
    return lib().fzo_payload_max_bytes(n)


def value_range(data):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    lo, hi = ctypes.c_float(), ctypes.c_float()
    rc = lib().fzo_range(_p(data), data.size, ctypes.byref(lo), ctypes.byref(hi))
    return rc, lo.value, hi.value


def eb_params(rng, eb_rel):
    eb_abs, r, e = ctypes.c_double(), ctypes.c_float(), ctypes.c_float()
    lib().fzo_eb_params(ctypes.c_float(rng), eb_rel, ctypes.byref(eb_abs), ctypes.byref(r), ctypes.byref(e))
    return eb_abs.value, np.float32(r.value), np.float32(e.value)


def dualquant(data, dims, ebx2_r, mode=MODE_FIXED, nthreads=1, length=None):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = data.size if length is None else length
    codes = np.zeros(n, dtype=np.uint16)
    ovf = lib().fzo_dualquant(_p(data), _p(codes), dims[0], dims[1], dims[2], float(ebx2_r), mode, nthreads)
    return codes, ovf


def bitshuffle_flags(words, ntiles, nthreads=1):
    """words: u32 array holding at least 1024*ntiles words.  Returns (out, bitflags[8T], byteflags[256T])."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    out = words.copy()
    bf = np.zeros(8 * ntiles, dtype=np.uint32)
    F = np.zeros(256 * ntiles, dtype=np.uint32)
    lib().fzo_bitshuffle_flags(_p(words), _p(out), ntiles, _p(bf), _p(F), nthreads)
    return out, bf, F


def exclusive_sum(a, nthreads=1):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros_like(a)
    lib().fzo_exclusive_sum(_p(a), _p(out), a.size, nthreads)
    return out


def encode(shuffled_bytes, presum, nblocks, nthreads=1):
    shuffled_bytes = np.ascontiguousarray(shuffled_bytes, dtype=np.uint8)
    presum = np.ascontiguousarray(presum, dtype=np.uint32)
    out = np.zeros(16 * int(presum[nblocks]) + 16, dtype=np.uint8)
    lib().fzo_encode(_p(shuffled_bytes), _p(out), _p(presum), nblocks, nthreads)
    return out[: 16 * int(presum[nblocks])]


def unshuffle(words, ntiles, nthreads=1):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    out = words.copy()
    lib().fzo_unshuffle(_p(words), _p(out), ntiles, nthreads)
    return out


def reconstruct(codes, dims, ebx2, nthreads=1):
    codes = np.ascontiguousarray(codes, dtype=np.uint16)
    n = dims[0] * dims[1] * dims[2]
    out = np.zeros(n, dtype=np.float32)
    lib().fzo_reconstruct(_p(codes), _p(out), dims[0], dims[1], dims[2], float(ebx2), nthreads)
    return out


def compress(data, dims, eb_rel, mode=MODE_FIXED, nthreads=1, debug=False):
    """Returns dict(status, info, payload bytes, tail u16, and with debug the intermediates)."""
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = data.size
    payload = np.zeros(payload_max_bytes(n) + 16, dtype=np.uint8)
    tail = np.zeros(2048, dtype=np.uint16)
    info = Info()
    dbg = {}
    if debug:
        W = flag_words(n)
        dbg = dict(codes=np.zeros(n + 2, dtype=np.uint16), shuffled=np.zeros((n + 3) // 2 + 2, dtype=np.uint32),
                   byteflags=np.zeros(32 * W, dtype=np.uint32), presum=np.zeros(32 * W + 1, dtype=np.uint32))
    rc = lib().fzo_compress(_p(data), n, dims[0], dims[1], dims[2], float(eb_rel), mode, nthreads, _p(payload),
                            _p(tail), ctypes.byref(info), _p(dbg.get("codes")), _p(dbg.get("shuffled")),
                            _p(dbg.get("byteflags")), _p(dbg.get("presum")))
    res = dict(status=rc, info=info.as_dict())
    if rc == 0:
        res["payload"] = payload[: info.payload_bytes].copy()
        res["tail"] = tail[: info.tail_len].copy() if mode == MODE_FIXED else np.zeros(0, np.uint16)
    if debug:
        dbg["codes"] = dbg["codes"][:n]
        res.update(dbg)
    return res


def decompress(payload, tail, n, dims, ebx2, nthreads=1, debug=False):
    payload = np.ascontiguousarray(payload, dtype=np.uint8)
    out = np.zeros(dims[0] * dims[1] * dims[2], dtype=np.float32)
    codes = np.zeros(n + 2, dtype=np.uint16) if debug else None
    t = None if tail is None else np.ascontiguousarray(tail, dtype=np.uint16)
    rc = lib().fzo_decompress(_p(payload), payload.size, _p(t), 0 if t is None else t.size, n, dims[0], dims[1],
                              dims[2], float(ebx2), nthreads, _p(out), _p(codes))
    if debug:
        return rc, out, codes[:n]
    return rc, out


def verify(xdata, odata):
    """verify_data (verify.hh:25-83) of reconstruction xdata against original odata."""
    x = np.ascontiguousarray(xdata, dtype=np.float32).ravel()
    o = np.ascontiguousarray(odata, dtype=np.float32).ravel()
    if x.size != o.size or x.size == 0:
        raise ValueError("verify needs two non-empty arrays of one length")
    st = Stats()
    lib().fzo_verify(_p(x), _p(o), x.size, ctypes.byref(st))
    return st.as_dict()


def compress_outlier(data, dims, eb_rel, side_cap=None, debug=False):
    """OUTLIER mode: dict(status, info, payload, tail, side (the side channel bytes))."""
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = data.size
    payload = np.zeros(payload_max_bytes(n) + 16, dtype=np.uint8)
    tail = np.zeros(2048, dtype=np.uint16)
    cap = lib().fzo_outlier_max_bytes(n) if side_cap is None else int(side_cap)
    side = np.zeros(max(cap, 16), dtype=np.uint8)
    codes = np.zeros(n + 2, dtype=np.uint16) if debug else None
    info = Info()
    rc = lib().fzo_compress_outlier(_p(data), n, dims[0], dims[1], dims[2], float(eb_rel), _p(payload), _p(tail),
                                    _p(side), cap, ctypes.byref(info), _p(codes))
    res = dict(status=rc, info=info.as_dict())
    if rc in (0, 7):
        res["payload"] = payload[: info.payload_bytes].copy()
        res["tail"] = tail[: info.tail_len].copy()
    if rc == 0:
        res["side"] = side[: info.side_bytes].copy()
    if debug:
        res["codes"] = codes[:n]
    return res


def decompress_outlier(payload, tail, side, n, dims, ebx2, vconst):
    payload = np.ascontiguousarray(payload, dtype=np.uint8)
    side = np.ascontiguousarray(side, dtype=np.uint8)
    out = np.zeros(dims[0] * dims[1] * dims[2], dtype=np.float32)
    t = None if tail is None else np.ascontiguousarray(tail, dtype=np.uint16)
    rc = lib().fzo_decompress_outlier(_p(payload), payload.size, _p(t), 0 if t is None else t.size, _p(side),
                                      side.size, n, dims[0], dims[1], dims[2], float(ebx2), float(vconst), _p(out))
    return rc, out


def side_lists(side):
    """(idx_d, val_d, idx_r, val_r) of a side channel byte string."""
    side = np.ascontiguousarray(side, dtype=np.uint8)
    nd, nr = (int(v) for v in side[:16].view(np.uint64))
    body = side[16:]
    w = body.view(np.uint32)
    return (w[:nd].copy(), w[nd:2 * nd].view(np.int32).copy(), w[2 * nd:2 * nd + nr].copy(),
            w[2 * nd + nr:2 * nd + 2 * nr].copy())
