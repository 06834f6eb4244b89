"""Python host binding of the gfx950 FZ library (lib/libfzgpu.so, C ABI in include/fz_capi.h).

PyTorch is used only for device memory and streams; every byte of the pipeline is computed
by the HIP kernels behind the C ABI.  There is no CPU fallback: when the library is missing
every call raises.

Mirrors of the reference's surface:
    run_szs(file, x, y, z, eb)          -> fz.cu:113 runSzs (prints fz.cu:207-209)
    lorenzo_construct(...)              -> lorenzo_var.cuh:520 launch_construct_LorenzoI_var
    bitshuffle_bitflag(...)             -> fz.cu:61 bitshuffleAndBitflag
    exclusive_sum(...)                  -> cub::DeviceScan::ExclusiveSum (fz.cu:188-192)
    encode_blocks(...)                  -> fz.cu:49 encodeDebug
    BatchPlan                           -> runSzs over many fields in one launch sequence
    verify(xdata, odata)                -> verify.hh:25 verify_data (quality metrics)
    quality_report(xdata, odata, ...)   -> quality_viewer.hh:97 echo_metric_gpu
"""
import ctypes
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libfzgpu.so")

MODE_FIXED = 0
MODE_LITERAL = 1
MODE_OUTLIER = 2
INVERSE_AUTO, INVERSE_WINDOWS, INVERSE_SPLIT = 0, 1, 2

STATUS = {0: "success", 1: "invalid argument", 2: "workspace too small", 3: "HIP runtime error",
          4: "constant field", 5: "non-finite value", 6: "corrupt payload", 7: "capacity exceeded",
          8: "not ready"}


class FzError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)} ({status})")


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_uint32), ("ndim", ctypes.c_uint32),
                ("vmin", ctypes.c_float), ("vmax", ctypes.c_float),
                ("ebx2_r", ctypes.c_float), ("ebx2", ctypes.c_float),
                ("eb_abs", ctypes.c_double),
                ("nblocks", ctypes.c_uint64), ("payload_bytes", ctypes.c_uint64),
                ("tail_len", ctypes.c_uint64), ("overflow", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


assert ctypes.sizeof(Result) == 64


class Sizes(ctypes.Structure):
    _fields_ = [("len", ctypes.c_uint64), ("tiles", ctypes.c_uint64), ("flag_words", ctypes.c_uint64),
                ("byte_flags", ctypes.c_uint64), ("payload_max_bytes", ctypes.c_uint64),
                ("tail_max_codes", ctypes.c_uint64), ("workspace_bytes", ctypes.c_uint64),
                ("side_max_bytes", ctypes.c_uint64)]


# every symbol declared in include/fz_capi.h
EXPORTS = ["fz_version", "fz_status_string", "fz_last_error", "fz_plan_create", "fz_plan_destroy", "fz_plan_sizes",
           "fz_workspace_init", "fz_compress_f32", "fz_compress_plain_f32", "fz_decompress_f32",
           "fz_compress_f32_sync", "fz_lorenzo_construct_f32", "fz_lorenzo_reconstruct_f32",
           "fz_bitshuffle_bitflag", "fz_bitunshuffle", "fz_exclusive_sum_u32", "fz_encode_blocks",
           "fz_plan_set_tile_layout", "fz_plan_set_inverse_form", "fz_plan_profile", "fz_plan_kernel_times",
           "fz_batch_plan_create", "fz_batch_plan_destroy", "fz_batch_plan_sizes", "fz_batch_field_sizes",
           "fz_batch_get_tuning", "fz_batch_set_tuning", "fz_batch_workspace_init", "fz_batch_compress_f32",
           "fz_batch_decompress_f32", "fz_batch_autotune", "fz_batch_profile", "fz_batch_kernel_times",
           "fz_batch_compress_graph", "fz_batch_decompress_graph", "fz_graph_launch", "fz_graph_destroy",
           "fz_verify_f32", "fz_decompress_verify_f32", "fz_compress_outlier_f32", "fz_decompress_outlier_f32",
           "fz_host_alloc", "fz_host_free", "fz_stager_create", "fz_stager_destroy", "fz_stager_sizes",
           "fz_stager_compress", "fz_stager_decompress", "fz_stager_wait", "fz_payload_check", "fz_side_check",
           "fz_compress_f32_noalloc", "fz_noalloc_scratch_bytes", "fz_batch_compress_outlier_f32",
           "fz_batch_decompress_outlier_f32"]


class _StatsRange(ctypes.Structure):
    _fields_ = [("min", ctypes.c_double), ("max", ctypes.c_double), ("rng", ctypes.c_double),
                ("std", ctypes.c_double)]


class _StatsReduced(ctypes.Structure):
    _fields_ = [("PSNR", ctypes.c_double), ("MSE", ctypes.c_double), ("NRMSE", ctypes.c_double),
                ("coeff", ctypes.c_double)]


class _StatsMaxErr(ctypes.Structure):
    _fields_ = [("abs", ctypes.c_double), ("rel", ctypes.c_double), ("pwrrel", ctypes.c_double),
                ("idx", ctypes.c_uint64)]


class _StatsAutocor(ctypes.Structure):
    _fields_ = [("lag_one", ctypes.c_double), ("lag_two", ctypes.c_double)]


class Stats(ctypes.Structure):
    """fz_stats = cusz_stats (type.h:185-194) with size_t as uint64."""
    _fields_ = [("odata", _StatsRange), ("xdata", _StatsRange), ("reduced", _StatsReduced),
                ("max_err", _StatsMaxErr), ("autocor", _StatsAutocor), ("user_eb", ctypes.c_double),
                ("len", ctypes.c_uint64)]

    def as_dict(self):
        def sub(st):
            return {k: getattr(st, k) for k, _ in st._fields_}
        return {"odata": sub(self.odata), "xdata": sub(self.xdata), "reduced": sub(self.reduced),
                "max_err": sub(self.max_err), "len": self.len}


assert ctypes.sizeof(Stats) == 160


class FieldShape(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("len", ctypes.c_uint64)]


class BatchSizes(ctypes.Structure):
    _fields_ = [("workspace_bytes", ctypes.c_uint64), ("total_len", ctypes.c_uint64), ("units", ctypes.c_uint64),
                ("nfields", ctypes.c_uint32), ("range_parts", ctypes.c_uint32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("range_chunk", ctypes.c_uint32), ("xcd_order", ctypes.c_uint32 * 3),
                ("gather_waves", ctypes.c_uint32), ("groups", ctypes.c_uint32), ("wave_tile", ctypes.c_uint32 * 3)]

    def as_dict(self):
        return {"range_chunk": self.range_chunk, "xcd_order": list(self.xcd_order),
                "gather_waves": self.gather_waves, "groups": self.groups,
                "wave_tile": list(self.wave_tile)}

_lib = None


def lib():
    """Load libfzgpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"FZ HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    # torch ships its own libamdhip64.so.7; load it first so that this process has a single
    # HIP runtime (the library's NEEDED entry then resolves to the copy already loaded)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32, f32, f64 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_double)
    L.fz_version.restype = ctypes.c_char_p
    L.fz_last_error.restype = ctypes.c_char_p
    L.fz_status_string.argtypes = [i32]
    L.fz_status_string.restype = ctypes.c_char_p
    L.fz_plan_create.argtypes = [ctypes.POINTER(vp), u32, u32, u32, u64, i32]
    L.fz_plan_destroy.argtypes = [vp]
    L.fz_plan_sizes.argtypes = [vp, ctypes.POINTER(Sizes)]
    L.fz_workspace_init.argtypes = [vp, vp, ctypes.c_size_t, vp]
    L.fz_compress_f32.argtypes = [vp, vp, f64, vp, ctypes.c_size_t, vp, vp, vp, vp]
    L.fz_compress_plain_f32.argtypes = [vp, vp, f64, vp, ctypes.c_size_t, vp, vp, vp, vp]
    L.fz_compress_f32_sync.argtypes = [vp, vp, f64, vp, ctypes.c_size_t, vp, vp, vp, ctypes.POINTER(Result), vp]
    L.fz_decompress_f32.argtypes = [vp, vp, vp, f32, vp, ctypes.c_size_t, vp, vp]
    L.fz_lorenzo_construct_f32.argtypes = [vp, vp, u32, u32, u32, u64, f64, i32, ctypes.POINTER(f32), vp]
    L.fz_lorenzo_reconstruct_f32.argtypes = [vp, vp, u32, u32, u32, f64, ctypes.POINTER(f32), vp]
    L.fz_bitshuffle_bitflag.argtypes = [vp, vp, u64, vp, vp, vp]
    L.fz_bitunshuffle.argtypes = [vp, vp, u64, vp]
    L.fz_exclusive_sum_u32.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), vp, vp, u64, vp]
    L.fz_encode_blocks.argtypes = [vp, vp, vp, u64, vp]
    L.fz_plan_profile.argtypes = [vp, i32]
    L.fz_plan_set_tile_layout.argtypes = [vp, i32]
    L.fz_plan_set_inverse_form.argtypes = [vp, i32]
    L.fz_plan_kernel_times.argtypes = [vp, ctypes.POINTER(f32), i32, ctypes.POINTER(i32)]
    L.fz_batch_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(FieldShape), u32, i32]
    L.fz_batch_plan_destroy.argtypes = [vp]
    L.fz_batch_plan_sizes.argtypes = [vp, ctypes.POINTER(BatchSizes)]
    L.fz_batch_field_sizes.argtypes = [vp, u32, ctypes.POINTER(Sizes)]
    L.fz_batch_get_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
    L.fz_batch_set_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
    L.fz_batch_workspace_init.argtypes = [vp, vp, ctypes.c_size_t, vp]
    L.fz_batch_compress_f32.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, vp]
    L.fz_batch_decompress_f32.argtypes = [vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp]
    L.fz_batch_autotune.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, i32, ctypes.POINTER(BatchTuning), vp]
    L.fz_batch_profile.argtypes = [vp, i32]
    L.fz_batch_kernel_times.argtypes = [vp, ctypes.POINTER(f32), i32, ctypes.POINTER(i32)]
    L.fz_batch_compress_graph.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, ctypes.POINTER(vp)]
    L.fz_batch_decompress_graph.argtypes = [vp, vp, vp, vp, vp, ctypes.c_size_t, vp, ctypes.POINTER(vp)]
    L.fz_graph_launch.argtypes = [vp, vp]
    L.fz_graph_destroy.argtypes = [vp]
    L.fz_verify_f32.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), vp, vp, vp, u64, vp]
    L.fz_compress_outlier_f32.argtypes = [vp, vp, f64, vp, ctypes.c_size_t, vp, vp, vp, u64, vp, vp]
    L.fz_host_alloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    L.fz_host_free.argtypes = [vp]
    L.fz_stager_create.argtypes = [ctypes.POINTER(vp), u32, u32, u32, u64, i32]
    L.fz_stager_destroy.argtypes = [vp]
    L.fz_stager_sizes.argtypes = [vp, ctypes.POINTER(Sizes)]
    L.fz_stager_compress.argtypes = [vp, vp, f64, ctypes.POINTER(u64)]
    L.fz_stager_decompress.argtypes = [vp, vp, u64, vp, f32, vp, ctypes.POINTER(u64)]
    L.fz_stager_wait.argtypes = [vp, u64, ctypes.POINTER(Result), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.fz_decompress_outlier_f32.argtypes = [vp, vp, vp, f32, f32, vp, u64, vp, ctypes.c_size_t, vp, vp]
    L.fz_decompress_verify_f32.argtypes = [vp, vp, vp, f32, vp, ctypes.c_size_t, vp, vp, vp, vp]
    L.fz_payload_check.argtypes = [vp, vp, u64, u64]
    L.fz_side_check.argtypes = [vp, vp, u64]
    L.fz_compress_f32_noalloc.argtypes = [vp, vp, f64, vp, vp, vp, vp]
    L.fz_noalloc_scratch_bytes.restype = u64
    L.fz_batch_compress_outlier_f32.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, vp, vp, vp]
    L.fz_batch_decompress_outlier_f32.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp]
    for name in EXPORTS:
        if name not in ("fz_version", "fz_status_string", "fz_last_error", "fz_noalloc_scratch_bytes"):
            getattr(L, name).restype = i32
    _lib = L
    return L


def version():
    return lib().fz_version().decode()


def _check(st, what):
    if st != 0:
        if st == 3:
            what = f"{what} [{lib().fz_last_error().decode()}]"
        raise FzError(st, what)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def _empty(nbytes, device="cuda"):
    import torch
    return torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device)


def _bucket(nbytes):
    """nbytes rounded up to one of eight sizes per octave above 1 MiB (at most 12.5% more).
    Batch workspaces and output arenas of a request stream come in as many sizes as there are
    batches; allocated at these sizes, a block the caching allocator freed is found again by
    the next batch of about the same size instead of splitting a larger one, so the allocator
    does not fragment its segments until it must free them all (a device-wide wait)."""
    n = max(int(nbytes), 16)
    if n <= 1 << 20:
        return n
    step = 1 << (n.bit_length() - 4)
    return (n + step - 1) // step * step


class Plan:
    """Geometry + mode of one field and its device workspace (allocated once, reused)."""

    def __init__(self, x, y, z, length=0, mode=MODE_FIXED, device="cuda", workspace=True):
        self.x, self.y, self.z = int(x), int(y), int(z)
        self.mode = mode
        self.device = device
        h = ctypes.c_void_p()
        _check(lib().fz_plan_create(ctypes.byref(h), self.x, self.y, self.z, int(length), mode), "fz_plan_create")
        self.handle = h
        s = Sizes()
        _check(lib().fz_plan_sizes(h, ctypes.byref(s)), "fz_plan_sizes")
        self.sizes = s
        self.len = s.len
        self.ws = None
        if workspace:
            self.ws = _empty(s.workspace_bytes, device)
            _check(lib().fz_workspace_init(h, _ptr(self.ws), self.ws.numel(), _stream()), "fz_workspace_init")

    @property
    def dims(self):
        return (self.x, self.y, self.z)

    def close(self):
        if self.handle is not None and _lib is not None:
            _lib.fz_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile_layout(self, wave_tile):
        """Tile-kernel layout of the fused compress: 1 = one wave per tile, 0 = one workgroup."""
        _check(lib().fz_plan_set_tile_layout(self.handle, int(wave_tile)), "fz_plan_set_tile_layout")

    def set_inverse_form(self, form):
        """Kernel form of decompress for 2D / 3D fields (same floats either way): INVERSE_AUTO,
        INVERSE_WINDOWS (decode into LDS, reconstruct straight to floats) or INVERSE_SPLIT
        (decode to codes, then reconstruct)."""
        _check(lib().fz_plan_set_inverse_form(self.handle, int(form)), "fz_plan_set_inverse_form")

    def profile(self, enable=True):
        _check(lib().fz_plan_profile(self.handle, 1 if enable else 0), "fz_plan_profile")

    def kernel_times(self):
        """Per-kernel milliseconds of the last compress/decompress call (waits for it)."""
        buf = (ctypes.c_float * 8)()
        cnt = ctypes.c_int(0)
        _check(lib().fz_plan_kernel_times(self.handle, buf, 8, ctypes.byref(cnt)), "fz_plan_kernel_times")
        return [buf[i] for i in range(cnt.value)]

    def alloc_outputs(self):
        import torch
        payload = _empty(self.sizes.payload_max_bytes + 16, self.device)
        tail = torch.empty(2048, dtype=torch.int16, device=self.device)
        res = torch.zeros(64, dtype=torch.uint8, device=self.device)
        return payload, tail, res

    def compress_async(self, d_in, eb_rel, payload, tail, res, fused=True, stream=None):
        """Launch only (graph-capturable); result record lands in `res` (64 device bytes)."""
        fn = lib().fz_compress_f32 if fused else lib().fz_compress_plain_f32
        _check(fn(self.handle, _ptr(d_in), float(eb_rel), _ptr(self.ws), self.ws.numel(), _ptr(payload), _ptr(tail),
                  _ptr(res), _stream(stream)), "fz_compress")

    def compress(self, d_in, eb_rel, fused=True, stream=None):
        """Returns (payload uint8 cuda tensor [payload_bytes], tail uint16 cuda tensor, result dict)."""
        import torch
        payload, tail, res = self.alloc_outputs()
        self.compress_async(d_in, eb_rel, payload, tail, res, fused, stream)
        r = read_result(res)
        if r["status"] != 0:
            raise FzError(r["status"], "compress")
        return payload[: r["payload_bytes"]], tail[: r["tail_len"]], r

    def check_payload(self, payload, tail_len):
        """fz_payload_check on host bytes (numpy uint8 array or bytes): raises FzError(6) when the
        payload's size, bitflags or the tail length disagree with this plan."""
        import numpy as np
        b = np.ascontiguousarray(np.frombuffer(payload, dtype=np.uint8) if isinstance(payload, (bytes, bytearray))
                                 else np.asarray(payload, dtype=np.uint8))
        _check(lib().fz_payload_check(self.handle, b.ctypes.data_as(ctypes.c_void_p) if b.size else None, b.size,
                                      int(tail_len)), "fz_payload_check")

    def check_side(self, side):
        """fz_side_check on a host side channel (OUTLIER plans)."""
        import numpy as np
        b = np.ascontiguousarray(np.asarray(side, dtype=np.uint8))
        _check(lib().fz_side_check(self.handle, b.ctypes.data_as(ctypes.c_void_p) if b.size else None, b.size),
               "fz_side_check")

    def decompress(self, payload, tail, ebx2, out=None, stream=None):
        import torch
        n = self.x * self.y * self.z
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        src = payload
        if payload.data_ptr() % 16:
            src = payload.clone()
        t = tail if (tail is not None and tail.numel() > 0) else None
        _check(lib().fz_decompress_f32(self.handle, _ptr(src), _ptr(t), float(ebx2), _ptr(self.ws), self.ws.numel(),
                                       _ptr(out), _stream(stream)), "fz_decompress")
        return out

    def compress_outlier(self, d_in, eb_rel, side_capacity=None, stream=None):
        """OUTLIER mode (plan made with MODE_OUTLIER): returns (payload, tail, result dict, side
        channel uint8 tensor [16 + 8 (n_delta + n_raw)])."""
        import torch
        payload, tail, res = self.alloc_outputs()
        cap = self.sizes.side_max_bytes if side_capacity is None else int(side_capacity)
        side = torch.zeros(max(cap, 16), dtype=torch.uint8, device=self.device)
        _check(lib().fz_compress_outlier_f32(self.handle, _ptr(d_in), float(eb_rel), _ptr(self.ws), self.ws.numel(),
                                             _ptr(payload), _ptr(tail), _ptr(side), cap, _ptr(res),
                                             _stream(stream)), "fz_compress_outlier")
        r = read_result(res)
        if r["status"] != 0:
            raise FzError(r["status"], "compress_outlier")
        nd, nr = (int(v) for v in side[:16].cpu().view(torch.int64).tolist())
        return payload[: r["payload_bytes"]], tail[: r["tail_len"]], r, side[: 16 + 8 * (nd + nr)]

    def decompress_outlier(self, payload, tail, side, ebx2, vconst, out=None, stream=None):
        import torch
        n = self.x * self.y * self.z
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        src = payload.clone() if payload.data_ptr() % 16 else payload
        sd = side.clone() if side.data_ptr() % 16 else side
        t = tail if (tail is not None and tail.numel() > 0) else None
        _check(lib().fz_decompress_outlier_f32(self.handle, _ptr(src), _ptr(t), float(ebx2), float(vconst), _ptr(sd),
                                               sd.numel(), _ptr(self.ws), self.ws.numel(), _ptr(out),
                                               _stream(stream)), "fz_decompress_outlier")
        return out

    def decompress_verify_async(self, payload, tail, ebx2, odata, out, d_stats, stream=None):
        """Launch only: decompress into out, then the metrics against odata into d_stats
        (160 device bytes).  payload must be 16-byte aligned."""
        t = tail if (tail is not None and tail.numel() > 0) else None
        _check(lib().fz_decompress_verify_f32(self.handle, _ptr(payload), _ptr(t), float(ebx2), _ptr(self.ws),
                                              self.ws.numel(), _ptr(out), _ptr(odata), _ptr(d_stats),
                                              _stream(stream)), "fz_decompress_verify")

    def decompress_verify(self, payload, tail, ebx2, odata, out=None, stream=None):
        """Decompress, then the quality metrics of the result against odata on the same
        stream.  Returns (out, stats dict)."""
        import torch
        n = self.x * self.y * self.z
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        src = payload
        if payload.data_ptr() % 16:
            src = payload.clone()
        d_stats = torch.zeros(ctypes.sizeof(Stats), dtype=torch.uint8, device=self.device)
        self.decompress_verify_async(src, tail, ebx2, odata, out, d_stats, stream)
        return out, read_stats(d_stats)


class HostBuffer:
    """Pinned host memory from fz_host_alloc, viewed as a numpy array (freed with the object)."""

    def __init__(self, n, dtype):
        import numpy as np
        self.dtype = np.dtype(dtype)
        self.nbytes = max(int(n) * self.dtype.itemsize, 16)
        p = ctypes.c_void_p()
        _check(lib().fz_host_alloc(ctypes.byref(p), self.nbytes), "fz_host_alloc")
        self.ptr = p
        buf = (ctypes.c_uint8 * self.nbytes).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=np.uint8)[: int(n) * self.dtype.itemsize].view(self.dtype)

    def close(self):
        if self.ptr is not None and _lib is not None:
            _lib.fz_host_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stager:
    """Host-resident requests of one geometry (fz_stager_*): depth requests in flight, the
    copy in, kernels and copy out of consecutive requests overlapped."""

    def __init__(self, x, y, z, length=0, depth=3):
        h = ctypes.c_void_p()
        _check(lib().fz_stager_create(ctypes.byref(h), int(x), int(y), int(z), int(length), int(depth)),
               "fz_stager_create")
        self.handle = h
        self.depth = depth
        s = Sizes()
        _check(lib().fz_stager_sizes(h, ctypes.byref(s)), "fz_stager_sizes")
        self.sizes = s
        self.keep = {}  # host buffers of requests in flight

    def compress(self, h_in, eb_rel):
        """Enqueue; h_in is a numpy float32 array (pinned, e.g. a HostBuffer's, for an
        asynchronous copy).  Returns the ticket."""
        t = ctypes.c_uint64()
        _check(lib().fz_stager_compress(self.handle, ctypes.c_void_p(h_in.ctypes.data), float(eb_rel),
                                        ctypes.byref(t)), "fz_stager_compress")
        self.keep[t.value] = h_in
        return t.value

    def decompress(self, h_payload, h_tail, ebx2, h_out):
        t = ctypes.c_uint64()
        tail_p = None if h_tail is None or h_tail.size == 0 else ctypes.c_void_p(h_tail.ctypes.data)
        _check(lib().fz_stager_decompress(self.handle, ctypes.c_void_p(h_payload.ctypes.data), h_payload.nbytes,
                                          tail_p, float(ebx2), ctypes.c_void_p(h_out.ctypes.data),
                                          ctypes.byref(t)), "fz_stager_decompress")
        self.keep[t.value] = (h_payload, h_tail, h_out)
        return t.value

    def wait(self, ticket):
        """Compress: (result dict, payload bytes copy, tail u16 copy); decompress: None."""
        import numpy as np
        r = Result()
        pp, tp = ctypes.c_void_p(), ctypes.c_void_p()
        st = lib().fz_stager_wait(self.handle, int(ticket), ctypes.byref(r), ctypes.byref(pp), ctypes.byref(tp))
        self.keep.pop(int(ticket), None)
        if st == 1:
            raise FzError(st, "fz_stager_wait")
        if not pp.value:
            _check(st, "fz_stager_wait")
            return None
        res = r.as_dict()
        payload = np.ctypeslib.as_array((ctypes.c_uint8 * max(int(r.payload_bytes), 1)).from_address(pp.value))
        tail = np.ctypeslib.as_array((ctypes.c_uint16 * max(int(r.tail_len), 1)).from_address(tp.value))
        return res, payload[: r.payload_bytes].copy(), tail[: r.tail_len].copy()

    def close(self):
        if self.handle is not None and _lib is not None:
            _lib.fz_stager_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_stats(stats_tensor):
    """fz_stats bytes on the device -> dict (waits for the stream)."""
    raw = stats_tensor.cpu().numpy().tobytes()
    return Stats.from_buffer_copy(raw[: ctypes.sizeof(Stats)]).as_dict()


def verify_async(d_xdata, d_odata, d_stats, temp=None, length=None, stream=None):
    """fz_verify_f32: metrics of d_xdata against d_odata into d_stats (160 device bytes);
    returns the temp tensor so a caller can reuse it."""
    n = d_odata.numel() if length is None else int(length)
    nbytes = ctypes.c_size_t(0)
    _check(lib().fz_verify_f32(None, ctypes.byref(nbytes), None, None, None, n, None), "fz_verify_f32(query)")
    if temp is None or temp.numel() < nbytes.value:
        temp = _empty(nbytes.value, d_odata.device)
    nbytes = ctypes.c_size_t(temp.numel())
    _check(lib().fz_verify_f32(_ptr(temp), ctypes.byref(nbytes), _ptr(d_stats), _ptr(d_xdata), _ptr(d_odata), n,
                               _stream(stream)), "fz_verify_f32")
    return temp


def verify(d_xdata, d_odata, stream=None):
    """verify_data (verify.hh:25-83) on the device: returns the stats dict."""
    import torch
    if d_xdata.numel() != d_odata.numel():
        raise FzError(1, "verify: arrays of different lengths")
    d_stats = torch.zeros(ctypes.sizeof(Stats), dtype=torch.uint8, device=d_odata.device)
    verify_async(d_xdata, d_odata, d_stats, stream=stream)
    return read_stats(d_stats)


def quality_report(d_xdata, d_odata, compressed_bytes=0, stream=None, echo=True):
    """echo_metric_gpu (quality_viewer.hh:97-110): cross metrics of the reconstruction, then
    the lag-1 and lag-2 autocorrelation of the original (odata vs odata shifted by 1 and 2),
    printed in the layout of print_metrics_cross / print_metrics_auto (:53-95)."""
    x = d_xdata.reshape(-1)
    o = d_odata.reshape(-1)
    s = verify(x, o, stream)
    n = o.numel()
    s["autocor"] = {"lag_one": verify(o[: n - 1], o[1:], stream)["reduced"]["coeff"] if n > 1 else float("nan"),
                    "lag_two": verify(o[: n - 2], o[2:], stream)["reduced"]["coeff"] if n > 2 else float("nan")}
    if echo:
        print(format_report(s, compressed_bytes))
    return s


def format_report(s, compressed_bytes=0):
    """Text of print_metrics_cross + print_metrics_auto (quality_viewer.hh:53-95) for float data."""
    def head(*c):
        return "  %-10s %16s %16s %16s %16s" % c

    def row(name, *v):
        return "  %-10s %16.8g %16.8g %16.8g %16.8g" % ((name,) + v)

    cr = (s["len"] * 4.0 / compressed_bytes) if compressed_bytes else float("inf")
    lines = ["", "quality metrics (using GPU checker):",
             head("", "data-len", "data-byte", "fp-type?", ""),
             "  %-10s %16d %16d %16s" % ("", s["len"], 4, "yes"),
             head("", "min", "max", "rng", "std"),
             row("origin", s["odata"]["min"], s["odata"]["max"], s["odata"]["rng"], s["odata"]["std"]),
             row("eb-lossy", s["xdata"]["min"], s["xdata"]["max"], s["xdata"]["rng"], s["xdata"]["std"]),
             head("", "abs-val", "abs-idx", "pw-rel", "VS-RNG"),
             row("max-error", s["max_err"]["abs"], s["max_err"]["idx"], s["max_err"]["pwrrel"], s["max_err"]["rel"]),
             head("", "CR", "NRMSE", "cross-cor", "PSNR"),
             row("metrics", cr, s["reduced"]["NRMSE"], s["reduced"]["coeff"], s["reduced"]["PSNR"])]
    if "autocor" in s:
        lines += [head("", "lag1-cor", "lag2-cor", "", ""),
                  "  %-10s %16f %16f" % ("auto", s["autocor"]["lag_one"], s["autocor"]["lag_two"])]
    return "\n".join(lines)


def read_result(res_tensor):
    """Copy the 64-byte device result record to the host (synchronises)."""
    raw = bytes(res_tensor[:64].cpu().numpy().tobytes())
    return Result.from_buffer_copy(raw).as_dict()


def read_results(res_tensor, n):
    """n consecutive 64-byte result records -> list of dicts (a device tensor is copied to the
    host, which synchronises; a host tensor is read as it is)."""
    raw = bytes(res_tensor[: 64 * n].cpu().numpy().tobytes())
    return [Result.from_buffer_copy(raw[64 * i: 64 * i + 64]).as_dict() for i in range(n)]


def _ptr_array(tensors):
    arr = (ctypes.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = None if t is None else t.data_ptr()
    return arr


def _al(n, a=256):
    return (int(n) + a - 1) // a * a


class Graph:
    """An instantiated hipGraph of one batched call on fixed buffers."""

    def __init__(self, handle, keep):
        self.handle = handle
        self._keep = keep  # the buffers the graph's kernels address

    def launch(self, stream=None):
        _check(lib().fz_graph_launch(self.handle, _stream(stream)), "fz_graph_launch")

    def close(self):
        if self.handle is not None and _lib is not None:
            _lib.fz_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchOutputs:
    """Device outputs of a batched compress: one payload arena and one tail arena with a
    256-byte-aligned slice per field, and the result records.  alloc(nbytes) may supply the
    arenas (a uint8 device tensor of at least nbytes, or None for a torch allocation)."""

    def __init__(self, plan, alloc=None):
        import torch

        def arena(nbytes):
            t = alloc(nbytes) if alloc is not None else None
            return t if t is not None else torch.empty(_bucket(nbytes), dtype=torch.uint8, device=plan.device)
        self.payload_off, self.tail_off = [], []
        po = to = 0
        for i in range(plan.nf):
            fs = plan.field_sizes(i)
            self.payload_off.append((po, fs.payload_max_bytes))
            self.tail_off.append((to, fs.tail_max_codes))
            po += _al(fs.payload_max_bytes + 16)
            to += _al(2 * fs.tail_max_codes + 2)
        self.payload_arena = arena(max(po, 256))
        self.tail_arena = arena(max(to, 256))
        self.payloads = [self.payload_arena[o: o + n + 16] for o, n in self.payload_off]
        self.tails = [self.tail_arena[o: o + 2 * n + 2].view(torch.int16) for o, n in self.tail_off]
        self.res = torch.zeros(64 * plan.nf, dtype=torch.uint8, device=plan.device)
        # OUTLIER mode: one side-channel arena, a slice of side_capacity(i) bytes per field
        self.sides, self.side_caps = [], []
        if plan.mode == MODE_OUTLIER:
            caps = [plan.side_capacity(i) for i in range(plan.nf)]
            offs, so = [], 0
            for c in caps:
                offs.append(so)
                so += _al(c)
            self.side_arena = arena(max(so, 256))
            self.side_arena.zero_()
            self.sides = [self.side_arena[o: o + c] for o, c in zip(offs, caps)]
            self.side_caps = caps

    def side_channels(self):
        """Each field's side channel (device uint8 views of 16 + 8 (n_delta + n_raw) bytes)."""
        import torch
        out = []
        for sd in self.sides:
            nd, nr = (int(v) for v in sd[:16].cpu().view(torch.int64).tolist())
            out.append(sd[: 16 + 8 * (nd + nr)])
        return out


class BatchPlan:
    """Many fields of uneven shapes (list of (x, y, z) or (x, y, z, len)) compressed by one
    launch sequence; each field's bytes equal those of Plan on that field alone."""

    def __init__(self, shapes, mode=MODE_FIXED, device="cuda", workspace=True, side_fraction=1.0):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.nf = len(self.shapes)
        self.mode = mode
        self.side_fraction = float(side_fraction)
        self.device = device
        arr = (FieldShape * max(self.nf, 1))()
        for i, s in enumerate(self.shapes):
            arr[i].x, arr[i].y, arr[i].z = s[0], s[1], s[2]
            arr[i].len = s[3] if len(s) > 3 else 0
        h = ctypes.c_void_p()
        _check(lib().fz_batch_plan_create(ctypes.byref(h), arr, self.nf, mode), "fz_batch_plan_create")
        self.handle = h
        bs = BatchSizes()
        _check(lib().fz_batch_plan_sizes(h, ctypes.byref(bs)), "fz_batch_plan_sizes")
        self.sizes = bs
        self.ws = None
        if workspace:
            self.attach_workspace(_empty(_bucket(bs.workspace_bytes), device))

    def attach_workspace(self, ws, stream=None):
        """Use the uint8 device tensor ws (at least sizes.workspace_bytes, 256-byte aligned) as
        the plan's workspace and put it into its reset state (fz_batch_workspace_init)."""
        if ws.numel() < self.sizes.workspace_bytes or ws.data_ptr() % 256:
            raise ValueError("workspace too small or not 256-byte aligned")
        self.ws = ws
        _check(lib().fz_batch_workspace_init(self.handle, _ptr(ws), ws.numel(), _stream(stream)),
               "fz_batch_workspace_init")

    def close(self):
        if self.handle is not None and _lib is not None:
            _lib.fz_batch_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field_sizes(self, i):
        s = Sizes()
        _check(lib().fz_batch_field_sizes(self.handle, i, ctypes.byref(s)), "fz_batch_field_sizes")
        return s

    def lens(self):
        return [self.field_sizes(i).len for i in range(self.nf)]

    @property
    def tuning(self):
        t = BatchTuning()
        _check(lib().fz_batch_get_tuning(self.handle, ctypes.byref(t)), "fz_batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, range_chunk=None, xcd_order=None, gather_waves=None, groups=None, wave_tile=None):
        t = BatchTuning()
        _check(lib().fz_batch_get_tuning(self.handle, ctypes.byref(t)), "fz_batch_get_tuning")
        if range_chunk is not None:
            t.range_chunk = range_chunk
        if xcd_order is not None:
            for i in range(3):
                t.xcd_order[i] = xcd_order[i]
        if gather_waves is not None:
            t.gather_waves = gather_waves
        if groups is not None:
            t.groups = groups
        if wave_tile is not None:
            for i in range(3):
                t.wave_tile[i] = wave_tile[i]
        _check(lib().fz_batch_set_tuning(self.handle, ctypes.byref(t)), "fz_batch_set_tuning")

    def alloc_outputs(self, alloc=None):
        return BatchOutputs(self, alloc)

    def _compress_args(self, inputs, ebs, out):
        assert len(inputs) == self.nf and len(ebs) == self.nf
        ins = _ptr_array(inputs)
        eb = (ctypes.c_double * self.nf)(*[float(e) for e in ebs])
        pays = _ptr_array(out.payloads)
        tails = _ptr_array(out.tails)
        return ins, eb, pays, tails

    def compress_async(self, inputs, ebs, out, stream=None):
        ins, eb, pays, tails = self._compress_args(inputs, ebs, out)
        _check(lib().fz_batch_compress_f32(self.handle, ins, eb, _ptr(self.ws), self.ws.numel(), pays, tails,
                                           _ptr(out.res), _stream(stream)), "fz_batch_compress_f32")

    def compress(self, inputs, ebs, stream=None):
        """Returns (outputs, list of result dicts); payload i = outputs.payloads[i][:payload_bytes]."""
        out = self.alloc_outputs()
        self.compress_async(inputs, ebs, out, stream)
        return out, read_results(out.res, self.nf)

    def compress_graph(self, inputs, ebs, out):
        ins, eb, pays, tails = self._compress_args(inputs, ebs, out)
        h = ctypes.c_void_p()
        _check(lib().fz_batch_compress_graph(self.handle, ins, eb, _ptr(self.ws), self.ws.numel(), pays, tails,
                                             _ptr(out.res), ctypes.byref(h)), "fz_batch_compress_graph")
        return Graph(h, (list(inputs), out, self.ws))

    def _decompress_args(self, payloads, tails, ebx2s, outs):
        assert len(payloads) == self.nf and len(ebx2s) == self.nf and len(outs) == self.nf
        pays = _ptr_array(payloads)
        tl = _ptr_array([t if (t is not None and t.numel() > 0) else None for t in tails]) if tails else None
        e = (ctypes.c_float * self.nf)(*[float(v) for v in ebx2s])
        return pays, tl, e, _ptr_array(outs)

    def decompress_async(self, payloads, tails, ebx2s, outs, stream=None):
        pays, tl, e, o = self._decompress_args(payloads, tails, ebx2s, outs)
        _check(lib().fz_batch_decompress_f32(self.handle, pays, tl, e, _ptr(self.ws), self.ws.numel(), o,
                                             _stream(stream)), "fz_batch_decompress_f32")

    def decompress(self, payloads, tails, ebx2s, stream=None):
        import torch
        outs = [torch.empty(max(s[0] * s[1] * s[2], 4), dtype=torch.float32, device=self.device)
                for s in self.shapes]
        self.decompress_async(payloads, tails, ebx2s, outs, stream)
        return [o[: s[0] * s[1] * s[2]] for o, s in zip(outs, self.shapes)]

    def side_capacity(self, i):
        """Side-channel bytes a batched OUTLIER compress reserves for field i: the largest
        channel the field can need, times the plan's side_fraction (1 by default)."""
        full = self.field_sizes(i).side_max_bytes
        return int(max(16, min(full, 16 + 8 * math.ceil((full - 16) / 8 * self.side_fraction))))

    def compress_outlier_async(self, inputs, ebs, out, stream=None):
        """OUTLIER-mode batch (plan made with MODE_OUTLIER): payloads, tails, side channels
        (out.sides) and records; launch only."""
        assert self.mode == MODE_OUTLIER
        ins, eb, pays, tails = self._compress_args(inputs, ebs, out)
        sides = _ptr_array(out.sides)
        caps = (ctypes.c_uint64 * self.nf)(*out.side_caps)
        _check(lib().fz_batch_compress_outlier_f32(self.handle, ins, eb, _ptr(self.ws), self.ws.numel(), pays, tails,
                                                   sides, caps, _ptr(out.res), _stream(stream)),
               "fz_batch_compress_outlier_f32")

    def compress_outlier(self, inputs, ebs, stream=None):
        out = self.alloc_outputs()
        self.compress_outlier_async(inputs, ebs, out, stream)
        return out, read_results(out.res, self.nf)

    def decompress_outlier_async(self, payloads, tails, sides, ebx2s, vconsts, outs, stream=None):
        assert self.mode == MODE_OUTLIER
        pays, tl, e, o = self._decompress_args(payloads, tails, ebx2s, outs)
        src = [sd.clone() if sd.data_ptr() % 16 else sd for sd in sides]
        sp = _ptr_array(src)
        sb = (ctypes.c_uint64 * self.nf)(*[int(sd.numel()) for sd in src])
        vc = (ctypes.c_float * self.nf)(*[float(v) for v in vconsts])
        _check(lib().fz_batch_decompress_outlier_f32(self.handle, pays, tl, e, vc, sp, sb, _ptr(self.ws),
                                                     self.ws.numel(), o, _stream(stream)),
               "fz_batch_decompress_outlier_f32")
        return src

    def decompress_outlier(self, payloads, tails, sides, ebx2s, vconsts, stream=None):
        import torch
        outs = [torch.empty(max(s[0] * s[1] * s[2], 4), dtype=torch.float32, device=self.device)
                for s in self.shapes]
        self.decompress_outlier_async(payloads, tails, sides, ebx2s, vconsts, outs, stream)
        return [o[: s[0] * s[1] * s[2]] for o, s in zip(outs, self.shapes)]

    def decompress_graph(self, payloads, tails, ebx2s, outs):
        pays, tl, e, o = self._decompress_args(payloads, tails, ebx2s, outs)
        h = ctypes.c_void_p()
        _check(lib().fz_batch_decompress_graph(self.handle, pays, tl, e, _ptr(self.ws), self.ws.numel(), o,
                                               ctypes.byref(h)), "fz_batch_decompress_graph")
        return Graph(h, (list(payloads), list(tails or []), list(outs), self.ws))

    def autotune(self, inputs, ebs, out, reps=3, stream=None):
        ins, eb, pays, tails = self._compress_args(inputs, ebs, out)
        t = BatchTuning()
        _check(lib().fz_batch_autotune(self.handle, ins, eb, _ptr(self.ws), self.ws.numel(), pays, tails,
                                       _ptr(out.res), int(reps), ctypes.byref(t), _stream(stream)),
               "fz_batch_autotune")
        return t.as_dict()

    def profile(self, enable=True):
        _check(lib().fz_batch_profile(self.handle, 1 if enable else 0), "fz_batch_profile")

    def kernel_times(self):
        buf = (ctypes.c_float * 8)()
        cnt = ctypes.c_int(0)
        _check(lib().fz_batch_kernel_times(self.handle, buf, 8, ctypes.byref(cnt)), "fz_batch_kernel_times")
        return [buf[i] for i in range(cnt.value)]


# ---------------------------------------------------------------- per-stage mirrors

def lorenzo_construct(d_data, dims, eb_abs, mode=MODE_FIXED, length=0, stream=None):
    import torch
    n = int(length) if length else dims[0] * dims[1] * dims[2]
    codes = torch.empty(n + 8, dtype=torch.int16, device=d_data.device)
    t = ctypes.c_float()
    _check(lib().fz_lorenzo_construct_f32(_ptr(d_data), _ptr(codes), dims[0], dims[1], dims[2], n, float(eb_abs),
                                          mode, ctypes.byref(t), _stream(stream)), "fz_lorenzo_construct_f32")
    return codes[:n], t.value


def lorenzo_reconstruct(d_codes, dims, eb_abs, stream=None):
    import torch
    n = dims[0] * dims[1] * dims[2]
    out = torch.empty(n, dtype=torch.float32, device=d_codes.device)
    t = ctypes.c_float()
    _check(lib().fz_lorenzo_reconstruct_f32(_ptr(d_codes), _ptr(out), dims[0], dims[1], dims[2], float(eb_abs),
                                            ctypes.byref(t), _stream(stream)), "fz_lorenzo_reconstruct_f32")
    return out, t.value


def bitshuffle_bitflag(d_words, length, stream=None):
    """d_words: int32 cuda tensor of >= 1024*floor(len/2048) words.  Returns (out, bitflags, byteflags)."""
    import torch
    T = int(length) // 2048
    out = d_words.clone()
    bf = torch.zeros(max(8 * T, 1), dtype=torch.int32, device=d_words.device)
    F = torch.zeros(max(256 * T, 1), dtype=torch.int32, device=d_words.device)
    _check(lib().fz_bitshuffle_bitflag(_ptr(d_words), _ptr(out), int(length), _ptr(bf), _ptr(F), _stream(stream)),
           "fz_bitshuffle_bitflag")
    return out, bf[: 8 * T], F[: 256 * T]


def bitunshuffle(d_words, length, stream=None):
    out = d_words.clone()
    _check(lib().fz_bitunshuffle(_ptr(d_words), _ptr(out), int(length), _stream(stream)), "fz_bitunshuffle")
    return out


def exclusive_sum(d_in, stream=None):
    """cub::DeviceScan::ExclusiveSum idiom: size query, temp allocation, run."""
    import torch
    n = d_in.numel()
    nbytes = ctypes.c_size_t(0)
    _check(lib().fz_exclusive_sum_u32(None, ctypes.byref(nbytes), _ptr(d_in), None, n, _stream(stream)),
           "fz_exclusive_sum_u32(query)")
    temp = _empty(nbytes.value, d_in.device)
    out = torch.empty_like(d_in)
    _check(lib().fz_exclusive_sum_u32(_ptr(temp), ctypes.byref(nbytes), _ptr(d_in), _ptr(out), n, _stream(stream)),
           "fz_exclusive_sum_u32")
    return out


def encode_blocks(d_in_bytes, d_presum, num_blocks, out_bytes, stream=None):
    import torch
    out = torch.zeros(max(int(out_bytes), 16), dtype=torch.uint8, device=d_in_bytes.device)
    _check(lib().fz_encode_blocks(_ptr(d_in_bytes), _ptr(out), _ptr(d_presum), int(num_blocks), _stream(stream)),
           "fz_encode_blocks")
    return out[: int(out_bytes)]


def run_szs(file_name, x, y, z, eb, mode=MODE_FIXED, fused=True):
    """fz.cu:113-209: read the file (float count from its size), compress on the GPU, print
    the three lines of fz.cu:207-209.  Returns (payload, tail, result)."""
    import numpy as np
    import torch
    file_size = os.path.getsize(file_name)
    n = file_size // 4
    host = np.fromfile(file_name, dtype=np.float32, count=n)
    d = torch.from_numpy(host).to("cuda")
    plan = Plan(x, y, z, n, mode)
    payload, tail, r = plan.compress(d, eb, fused=fused)
    csize = r["payload_bytes"]
    print(f"original size: {file_size}")
    print(f"compressed size: {csize}")
    print("compression ratio: %f" % (np.float32(file_size) / np.float32(csize)))
    return payload, tail, r
