"""GPU parity of the outlier mode (fz_compress_outlier_f32 / fz_decompress_outlier_f32) against
the oracle's restatement: payload, tail, side channel and result record byte for byte, the
inverse bit for bit, on fields with large deltas, non-finite values, oversized values and
constant fields; full-size round trips through error-bound and bit-identity properties."""
import ctypes
import sys

import numpy as np
import pytest

import fzgpu
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _np(t):
    return t.cpu().numpy()


def _field(dims, seed, spikes=0, nans=0, infs=0):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    x = np.cumsum(rng.normal(size=n)).astype(np.float32)
    if spikes:
        x[rng.integers(0, n, spikes)] = np.float32(1e7)
    if nans:
        x[rng.integers(0, n, nans)] = np.nan
    if infs:
        x[rng.integers(0, n, infs)] = np.float32(np.inf)
        x[rng.integers(0, n, infs)] = np.float32(-np.inf)
    return x


def _check(x, dims, eb, length=None):
    n = x.size
    say(f"case outlier dims={dims} len={n} eb={eb}")
    ref = O.compress_outlier(x, dims, eb)
    assert ref["status"] == 0
    plan = fzgpu.Plan(*dims, length=n, mode=fzgpu.MODE_OUTLIER)
    d_x = _to_dev(x)
    payload, tail, r, side = plan.compress_outlier(d_x, eb)
    info = ref["info"]
    for k_gpu, k_ref in [("payload_bytes", "payload_bytes"), ("nblocks", "K"), ("overflow", "overflow"),
                         ("vmin", "vmin"), ("vmax", "vmax"), ("ebx2", "ebx2"), ("ebx2_r", "ebx2_r"),
                         ("eb_abs", "eb_abs"), ("tail_len", "tail_len")]:
        assert r[k_gpu] == info[k_ref], (k_gpu, r[k_gpu], info[k_ref])
    assert np.array_equal(_np(payload), ref["payload"])
    assert np.array_equal(_np(tail).view(np.uint16), ref["tail"])
    assert np.array_equal(_np(side), ref["side"])
    out = plan.decompress_outlier(payload, tail, side, r["ebx2"], r["vmin"])
    rc, ref_out = O.decompress_outlier(ref["payload"], ref["tail"], ref["side"], n, dims, info["ebx2"], info["vmin"])
    assert rc == 0
    assert np.array_equal(_np(out).view(np.uint32), ref_out.view(np.uint32))
    return r, ref, _np(out)


SHAPES = [(5000, 1, 1), (4096, 1, 1), (2047, 1, 1), (100, 77, 1), (16, 16, 1), (17, 33, 1), (3600, 20, 1),
          (33, 20, 17), (64, 64, 64), (7, 1, 900), (45, 46, 47), (9, 50, 40), (1030, 9, 11)]


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("eb", [1e-6, 1e-3])
def test_outlier_mode_matches_oracle(dims, eb):
    x = _field(dims, sum(dims), spikes=5, nans=4, infs=2)
    _check(x, dims, eb)


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("eb", [1e-4, 1e-5])
def test_finite_fields_with_spikes(dims, eb):
    # all finite: the fixed mode's float paths run and only the tiles with a spike's
    # overflowing delta are redone on the integer path
    x = _field(dims, 3 * sum(dims), spikes=5)
    r, ref, _ = _check(x, dims, eb)
    assert ref["info"]["n_raw"] == 0


@pytest.mark.parametrize("dims", [(5000, 1, 1), (100, 77, 1), (64, 64, 64)])
def test_fields_needing_no_side_channel_take_the_fixed_payload(dims):
    x = _field(dims, 9)
    r, ref, _ = _check(x, dims, 1e-2)
    assert r["overflow"] == 0 and ref["info"]["n_raw"] == 0
    fixed = fzgpu.Plan(*dims)
    payload, tail, _ = fixed.compress(_to_dev(x), 1e-2)
    assert np.array_equal(_np(payload), ref["payload"])


def test_len_beyond_xyz_and_ragged_tail():
    dims = (1000, 3, 1)
    x = _field((6000, 1, 1), 5, spikes=3, nans=2)
    _check(x, dims, 1e-6)


def test_constant_fields_all_non_finite_and_oversized_values():
    c = np.full(12000, -3.25, np.float32)
    r, _, out = _check(c, (300, 40, 1), 1e-3)
    assert r["ebx2"] == 0.0 and np.array_equal(out, c)
    c[[5, 7000, 11999]] = [np.nan, np.inf, np.nan]
    _, _, out = _check(c, (300, 40, 1), 1e-3)
    assert np.array_equal(out.view(np.uint32), c.view(np.uint32))
    a = np.full(5000, np.nan, np.float32)
    _check(a, (5000, 1, 1), 1e-3)
    rng = np.random.default_rng(3)
    big = (np.float32(1e9) + rng.integers(0, 64, 4096).astype(np.float32) * np.float32(64)).astype(np.float32)
    r, ref, out = _check(big, (4096, 1, 1), 1e-3)
    assert ref["info"]["n_raw"] == big.size and np.array_equal(out, big)


def test_capacity_status_and_mode_checks():
    import torch
    dims = (100, 77, 1)
    x = _field(dims, 4, spikes=4, nans=3)
    say("case outlier capacity")
    ref = O.compress_outlier(x, dims, 1e-6)
    need = ref["info"]["side_bytes"]
    plan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
    with pytest.raises(fzgpu.FzError) as e:
        plan.compress_outlier(_to_dev(x), 1e-6, side_capacity=need - 8)
    assert e.value.status == 7
    _, _, r, side = plan.compress_outlier(_to_dev(x), 1e-6, side_capacity=need)
    assert np.array_equal(_np(side), ref["side"])
    # the other entry points refuse an outlier plan, and the outlier ones a FIXED plan
    L = fzgpu.lib()
    payload, tail, res = plan.alloc_outputs()
    d = _to_dev(x)
    st = L.fz_compress_f32(plan.handle, ctypes.c_void_p(d.data_ptr()), 1e-3, ctypes.c_void_p(plan.ws.data_ptr()),
                           plan.ws.numel(), ctypes.c_void_p(payload.data_ptr()), ctypes.c_void_p(tail.data_ptr()),
                           ctypes.c_void_p(res.data_ptr()), None)
    assert st == 1
    fixed = fzgpu.Plan(*dims)
    side_buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    st = L.fz_compress_outlier_f32(fixed.handle, ctypes.c_void_p(d.data_ptr()), 1e-3,
                                   ctypes.c_void_p(fixed.ws.data_ptr()), fixed.ws.numel(),
                                   ctypes.c_void_p(payload.data_ptr()), ctypes.c_void_p(tail.data_ptr()),
                                   ctypes.c_void_p(side_buf.data_ptr()), 64, ctypes.c_void_p(res.data_ptr()), None)
    assert st == 1


def test_range_overflow_is_reported():
    x = np.linspace(-3e38, 3e38, 5000).astype(np.float32)
    say("case outlier range overflow")
    ref = O.compress_outlier(x, (5000, 1, 1), 1e-3)
    assert ref["status"] == 5
    plan = fzgpu.Plan(5000, 1, 1, mode=fzgpu.MODE_OUTLIER)
    with pytest.raises(fzgpu.FzError) as e:
        plan.compress_outlier(_to_dev(x), 1e-3)
    assert e.value.status == 5


@pytest.mark.parametrize("dims", [(3600, 1800, 1), (500, 500, 100)])
def test_full_size_round_trip_with_planted_values(dims):
    import torch
    n = dims[0] * dims[1] * dims[2]
    say(f"case outlier full size {dims}")
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(n, device="cuda", generator=g).cumsum(0)
    idx = torch.randint(0, n, (2000,), device="cuda", generator=g)
    x[idx[:500]] = float("nan")
    x[idx[500:1000]] = float("inf")
    x[idx[1000:]] = 1e7
    plan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
    payload, tail, r, side = plan.compress_outlier(x, 1e-5)
    out = plan.decompress_outlier(payload, tail, side, r["ebx2"], r["vmin"])
    fin = torch.isfinite(x)
    err = (out[fin].double() - x[fin].double()).abs().max().item()
    assert err <= r["eb_abs"] + 4 * 2.0 ** -24 * x[fin].abs().max().item()
    assert torch.equal(out[~fin].view(torch.int32), x[~fin].view(torch.int32))
    nd, nr = (int(v) for v in side[:16].cpu().view(torch.int64).tolist())
    assert nd == r["overflow"] and nr == int((~fin).sum().item()) and nd > 0
    # the payload is deterministic: a second call gives the same bytes
    p2, t2, r2, s2 = plan.compress_outlier(x, 1e-5)
    assert torch.equal(p2, payload) and torch.equal(s2, side) and torch.equal(t2, tail)


def test_maximum_length_field_with_planted_values():
    """The largest field a plan accepts (len = 2^32 - 4196, 17.2 GB, over 2^20 tiles) in the
    outlier mode, with NaN, +-Inf and spikes planted up to the last positions: finite values
    within the bound, the planted values bit for bit, the channel's counts as planted."""
    import torch
    n = 0xFFFFF000 - 100
    torch.cuda.empty_cache()
    say(f"case outlier maximum length n={n}")
    x = torch.arange(n, dtype=torch.float32, device="cuda")
    x.mul_(1e-3).sin_()
    pos = torch.tensor([0, 5, 1 << 31, (1 << 32) - 9000, n - 2049, n - 2, n - 1], dtype=torch.int64, device="cuda")
    vals = torch.tensor([float("nan"), float("inf"), 1e7, -float("inf"), float("nan"), 1e7, float("nan")],
                        device="cuda")
    x[pos] = vals
    plan = fzgpu.Plan(n, 1, 1, mode=fzgpu.MODE_OUTLIER)
    # eb 1e-6 of a 1e7 range: the spikes' deltas overflow the 16-bit codes
    payload, tail, r, side = plan.compress_outlier(x, 1e-6, side_capacity=1 << 20)
    nd, nr = (int(v) for v in side[:16].cpu().view(torch.int64).tolist())
    assert nr == 5 and nd == r["overflow"] and nd > 0
    out = plan.decompress_outlier(payload, tail, side, r["ebx2"], r["vmin"])
    fin = torch.isfinite(x)
    assert torch.equal(out[pos].view(torch.int32)[~torch.isfinite(vals)], vals.view(torch.int32)[~torch.isfinite(vals)])
    err = float(out.sub_(x).masked_fill_(~fin, 0).abs_().max())
    assert err <= r["eb_abs"] + 4 * 2.0**-24 * 1e7
    del x, out, payload, tail, side, plan
    torch.cuda.empty_cache()
