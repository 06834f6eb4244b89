"""GPU parity of the quality metrics (fz_verify_f32, fz_decompress_verify_f32) against the
oracle's verify_data restatement (verify.hh:25-83).

Bar: extrema, ranges, the largest |x-o|, its index, its ratio to the range and the largest
pointwise ratio are exact (the GPU keeps the reference's float quantities and tie rules);
sums-derived values (std, MSE, NRMSE, correlation) agree to a relative 1e-9 and PSNR to an
absolute 1e-9 dB, since the GPU adds in another order and around another centre.
"""
import math
import sys

import numpy as np
import pytest

import fzgpu
import oracle as O

pytestmark = pytest.mark.gpu

EXACT = [("odata", "min"), ("odata", "max"), ("odata", "rng"), ("xdata", "min"), ("xdata", "max"),
         ("xdata", "rng"), ("max_err", "abs"), ("max_err", "idx"), ("max_err", "rel"), ("max_err", "pwrrel")]
REL = [("odata", "std"), ("xdata", "std"), ("reduced", "MSE"), ("reduced", "NRMSE"), ("reduced", "coeff")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _same(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


def _close(a, b, rel=1e-9, abs_=0.0):
    if math.isnan(a) or math.isnan(b) or math.isinf(a) or math.isinf(b):
        return _same(a, b)
    return abs(a - b) <= max(rel * abs(b), abs_)


def _compare(g, r):
    assert g["len"] == r["len"]
    for a, b in EXACT:
        assert _same(float(g[a][b]), float(r[a][b])), (a, b, g[a][b], r[a][b])
    for a, b in REL:
        assert _close(g[a][b], r[a][b]), (a, b, g[a][b], r[a][b])
    assert _close(g["reduced"]["PSNR"], r["reduced"]["PSNR"], rel=0.0, abs_=1e-9), (g["reduced"], r["reduced"])


def _pair(n, seed, eb=1e-3, zero=0.05, offset=0.0):
    rng = np.random.default_rng(seed)
    o = (np.cumsum(rng.normal(size=n)) + offset).astype(np.float32)
    o[rng.random(n) < zero] = 0
    r = float(o.max() - o.min()) or 1.0
    x = (o + rng.uniform(-eb * r, eb * r, n).astype(np.float32)).astype(np.float32)
    return x, o


@pytest.mark.parametrize("n", [1, 2, 3, 100, 1023, 2047, 2048, 2049, 4095, 100003, 1 << 20, 3600 * 1800])
def test_verify_matches_oracle(n):
    say(f"case verify n={n}")
    x, o = _pair(n, n)
    _compare(fzgpu.verify(_to_dev(x), _to_dev(o)), O.verify(x, o))


@pytest.mark.parametrize("offset", [1e4, -3e5])
def test_verify_far_from_zero(offset):
    # values far from 0 with a small spread: the centred sums must not cancel
    n = 500000
    say(f"case verify offset={offset}")
    x, o = _pair(n, 7, offset=offset, zero=0.0)
    _compare(fzgpu.verify(_to_dev(x), _to_dev(o)), O.verify(x, o))


@pytest.mark.parametrize("lag", [1, 2, 3])
def test_unaligned_views_as_in_the_autocorrelation_calls(lag):
    # quality_viewer.hh:105-107: verify(origin, origin + lag, len - lag)
    n = 300001
    say(f"case verify lag={lag}")
    _, o = _pair(n, 11)
    d = _to_dev(o)
    g = fzgpu.verify(d[: n - lag], d[lag:])
    _compare(g, O.verify(o[: n - lag], o[lag:]))


def test_ties_of_the_largest_error_keep_the_first_index():
    # equal largest errors in different chunks, workgroups and lanes: the first one wins
    n = 5_000_000
    say("case verify ties")
    o = np.linspace(1, 2, n).astype(np.float32)
    x = o.copy()
    planted = (4_999_000, 3_000_001, 2_100_000, 777_777, 1_500_000, 777_778)
    for i in planted:
        o[i], x[i] = 1.5, 1.75
    g = fzgpu.verify(_to_dev(x), _to_dev(o))
    _compare(g, O.verify(x, o))
    assert (g["max_err"]["abs"], g["max_err"]["idx"]) == (0.25, min(planted))


def test_constant_field_and_identical_arrays():
    say("case verify constant / identical")
    c = np.full(10000, 2.5, np.float32)
    x = c + np.float32(0.5)
    _compare(fzgpu.verify(_to_dev(x), _to_dev(c)), O.verify(x, c))
    o = np.linspace(0, 1, 70000).astype(np.float32)
    _compare(fzgpu.verify(_to_dev(o), _to_dev(o)), O.verify(o, o))


def test_non_numbers_follow_the_reference_scan():
    say("case verify nan")
    x, o = _pair(10000, 3)
    x2 = x.copy()
    x2[5000] = np.nan  # skipped by every comparison, poisons the sums
    _compare(fzgpu.verify(_to_dev(x2), _to_dev(o)), O.verify(x2, o))
    x3 = x.copy()
    x3[0] = np.nan  # element 0 starts the scans: never replaced
    _compare(fzgpu.verify(_to_dev(x3), _to_dev(o)), O.verify(x3, o))


def test_verify_rejects_bad_arguments():
    import torch
    d = torch.zeros(16, dtype=torch.float32, device="cuda")
    L = fzgpu.lib()
    import ctypes
    nb = ctypes.c_size_t(0)
    assert L.fz_verify_f32(None, ctypes.byref(nb), None, None, None, 0, None) == 1
    assert L.fz_verify_f32(None, ctypes.byref(nb), None, None, None, 16, None) == 0
    temp = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    small = ctypes.c_size_t(nb.value - 1)
    st = torch.zeros(160, dtype=torch.uint8, device="cuda")
    assert L.fz_verify_f32(ctypes.c_void_p(temp.data_ptr()), ctypes.byref(small), ctypes.c_void_p(st.data_ptr()),
                           ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(d.data_ptr()), 16, None) == 2
    # 2-byte offset: not even float-aligned
    assert L.fz_verify_f32(ctypes.c_void_p(temp.data_ptr()), ctypes.byref(nb), ctypes.c_void_p(st.data_ptr()),
                           ctypes.c_void_p(d.data_ptr() + 2), ctypes.c_void_p(d.data_ptr()), 8, None) == 1


@pytest.mark.parametrize("dims,eb", [((3600, 1800, 1), 1e-3), ((64, 64, 64), 1e-4), ((5000, 1, 1), 1e-2),
                                     ((100, 77, 1), 1e-3)])
def test_decompress_verify_matches_oracle(dims, eb):
    say(f"case decompress_verify dims={dims} eb={eb}")
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(5)
    o = np.cumsum(rng.normal(size=n)).astype(np.float32)
    plan = fzgpu.Plan(*dims)
    d_o = _to_dev(o)
    payload, tail, r = plan.compress(d_o, eb)
    out, g = plan.decompress_verify(payload, tail, r["ebx2"], d_o)
    x = out.cpu().numpy()
    ref = O.verify(x, o)
    _compare(g, ref)
    # the same numbers as a separate verify call, and the compressor's error bound holds
    assert g == fzgpu.verify(out, d_o)
    assert g["max_err"]["abs"] <= r["eb_abs"] * (1 + 1e-6) + 4 * 2.0 ** -24 * float(np.abs(o).max())


def test_single_field_entry_points_replay_in_a_graph():
    """fz_compress_f32, fz_decompress_verify_f32 and the outlier pair launch kernels only, so
    they capture into a graph; every replay (with eager work and syncs between replays) gives
    the eager bytes and numbers."""
    import ctypes

    import torch
    dims = (3600, 180, 1)
    n = dims[0] * dims[1]
    say("case graph capture of the single-field entry points")
    x, o = _pair(n, 21)
    o[::997] = np.nan  # the outlier plan keeps these verbatim
    d_o = _to_dev(np.nan_to_num(o, nan=0.0))
    d_on = _to_dev(o)
    plan = fzgpu.Plan(*dims)
    payload, tail, res = plan.alloc_outputs()
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    st = torch.zeros(160, dtype=torch.uint8, device="cuda")
    oplan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
    opay, otail, ores = oplan.alloc_outputs()
    cap = oplan.sizes.side_max_bytes
    side = torch.zeros(cap, dtype=torch.uint8, device="cuda")
    oout = torch.empty(n, dtype=torch.float32, device="cuda")
    L = fzgpu.lib()
    P = fzgpu._ptr

    # eager pass fixes ebx2 / vmin / side size, the values the captured calls are given
    plan.compress_async(d_o, 1e-3, payload, tail, res)
    r = fzgpu.read_result(res)
    L.fz_compress_outlier_f32(oplan.handle, P(d_on), 1e-3, P(oplan.ws), oplan.ws.numel(), P(opay), P(otail), P(side),
                              cap, P(ores), fzgpu._stream())
    orr = fzgpu.read_result(ores)
    nd, nr = (int(v) for v in side[:16].cpu().view(torch.int64).tolist())
    sbytes = 16 + 8 * (nd + nr)
    assert nr == int(np.isnan(o).sum())

    def calls():
        plan.compress_async(d_o, 1e-3, payload, tail, res)
        plan.decompress_verify_async(payload, tail, r["ebx2"], d_o, out, st)
        fzgpu._check(L.fz_compress_outlier_f32(oplan.handle, P(d_on), 1e-3, P(oplan.ws), oplan.ws.numel(), P(opay),
                                               P(otail), P(side), cap, P(ores), fzgpu._stream()), "outlier")
        fzgpu._check(L.fz_decompress_outlier_f32(oplan.handle, P(opay), P(otail), float(orr["ebx2"]),
                                                 float(orr["vmin"]), P(side), sbytes, P(oplan.ws), oplan.ws.numel(),
                                                 P(oout), fzgpu._stream()), "outlier inverse")

    calls()
    torch.cuda.synchronize()
    pb, ob, tl = r["payload_bytes"], orr["payload_bytes"], r["tail_len"]

    def results():  # the bytes the calls define (buffers are larger)
        return (payload[:pb], tail[:tl], res, out, st, opay[:ob], side[:sbytes], oout)

    ref = [t.clone() for t in results()]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        calls()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        calls()
    for rep in range(4):
        for t in (payload, out, st, opay, side, oout):
            t.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(results(), ref):
            assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), rep
