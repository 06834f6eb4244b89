"""OUTLIER-mode batches (fz_batch_compress_outlier_f32 / fz_batch_decompress_outlier_f32): a
batch mixing fields the FIXED mode cannot encode (deltas beyond 32767, constant fields, NaN
and +-Inf, values too large to quantise) with ordinary ones, every field's payload, tail, side
channel and record against the oracle's outlier mode and the single-field entry point, byte
for byte, and the batched inverse bit for bit; two calls in a row on one shared workspace."""
import sys

import numpy as np
import pytest

import fzgpu
import oracle as O
from fzgpu import datagen as D

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.cpu().numpy()


def _mix():
    """(dims, host field, eb) of a mixed batch."""
    rng = np.random.default_rng(7)
    out = []
    shapes = [(5000, 1, 1), (100, 77, 1), (33, 20, 17), (2047, 1, 1), (3600, 20, 1), (64, 64, 64), (7, 1, 900),
              (45, 46, 47), (1030, 9, 11), (70000, 1, 1), (300, 301, 1), (16, 16, 1)]
    for i, d in enumerate(shapes):
        n = d[0] * d[1] * d[2]
        x = np.cumsum(rng.normal(size=n)).astype(np.float32)
        kind = i % 6
        if kind == 1:    # deltas beyond the u16 code (small eb, spikes)
            x[rng.integers(0, n, 5)] = np.float32(1e7)
        elif kind == 2:  # constant field
            x[:] = np.float32(3.25)
        elif kind == 3:  # NaN and +-Inf
            x[rng.integers(0, n, 3)] = np.nan
            x[rng.integers(0, n, 2)] = np.inf
            x[rng.integers(0, n, 2)] = -np.inf
        elif kind == 4:  # values too large to quantise exactly (|x r| >= 2^24)
            x = (x + np.float32(1e9)).astype(np.float32)
        eb = (1e-6, 1e-3, 1e-4)[i % 3]
        out.append((d, x, eb))
    return out


def test_outlier_batch_matches_oracle_and_single_field():
    import torch
    mix = _mix()
    bp = fzgpu.BatchPlan([d for d, _, _ in mix], mode=fzgpu.MODE_OUTLIER)
    d_in = [torch.from_numpy(x).cuda() for _, x, _ in mix]
    ebs = [eb for _, _, eb in mix]
    say(f"outlier batch of {len(mix)} fields")
    runs = []
    for rep in range(2):  # the second call reuses the workspace the first left behind
        out, res = bp.compress_outlier(d_in, ebs)
        torch.cuda.synchronize()
        runs.append((out, res))
    (out, res), (out2, res2) = runs
    assert res == res2
    sides = out.side_channels()
    for i, (dims, x, eb) in enumerate(mix):
        say(f"  field {i} dims={dims} eb={eb}")
        ref = O.compress_outlier(x, dims, eb)
        assert ref["status"] == 0 and res[i]["status"] == 0, (i, ref["status"], res[i]["status"])
        info = ref["info"]
        for k_gpu, k_ref in [("payload_bytes", "payload_bytes"), ("nblocks", "K"), ("overflow", "overflow"),
                             ("vmin", "vmin"), ("vmax", "vmax"), ("ebx2", "ebx2"), ("ebx2_r", "ebx2_r"),
                             ("eb_abs", "eb_abs"), ("tail_len", "tail_len")]:
            assert res[i][k_gpu] == info[k_ref], (i, k_gpu, res[i][k_gpu], info[k_ref])
        pay = _np(out.payloads[i][: res[i]["payload_bytes"]])
        assert np.array_equal(pay, ref["payload"]), i
        assert np.array_equal(_np(out.tails[i][: res[i]["tail_len"]]).view(np.uint16), ref["tail"]), i
        assert np.array_equal(_np(sides[i]), ref["side"]), i
        assert torch.equal(out2.payloads[i][: res[i]["payload_bytes"]], out.payloads[i][: res[i]["payload_bytes"]])
        # the single-field entry point gives the same bytes
        plan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
        p1, t1, r1, s1 = plan.compress_outlier(d_in[i], eb)
        assert r1 == res[i] and np.array_equal(_np(p1), pay) and np.array_equal(_np(s1), _np(sides[i])), i
    # batched inverse, twice, against the oracle's
    pays = [out.payloads[i][: res[i]["payload_bytes"]] for i in range(len(mix))]
    tails = [out.tails[i][: res[i]["tail_len"]] for i in range(len(mix))]
    for rep in range(2):
        xs = bp.decompress_outlier(pays, tails, sides, [r["ebx2"] for r in res], [r["vmin"] for r in res])
        torch.cuda.synchronize()
        for i, (dims, x, eb) in enumerate(mix):
            ref = O.compress_outlier(x, dims, eb)
            rc, xo = O.decompress_outlier(ref["payload"], ref["tail"], ref["side"], x.size, dims, ref["info"]["ebx2"],
                                          ref["info"]["vmin"])
            assert rc == 0
            assert np.array_equal(_np(xs[i]).view(np.uint32), xo.view(np.uint32)), (rep, i)


def test_fixed_batch_entry_points_refuse_an_outlier_plan_and_back():
    import torch
    dims = [(5000, 1, 1), (100, 77, 1)]
    bo = fzgpu.BatchPlan(dims, mode=fzgpu.MODE_OUTLIER)
    d = [torch.ones(5000, device="cuda"), torch.ones(7700, device="cuda")]
    with pytest.raises(fzgpu.FzError):
        bo.compress_async(d, [1e-3, 1e-3], bo.alloc_outputs())
    bf = fzgpu.BatchPlan(dims)
    with pytest.raises(AssertionError):
        bf.compress_outlier(d, [1e-3, 1e-3])


def test_outlier_batch_of_generated_fields_round_trip():
    """Sixteen generated fields (1D / 2D / 3D, up to 4M floats) with a few NaN and spike values
    planted: the batched inverse restores every finite value within the bound and every
    verbatim value bit for bit."""
    import torch
    spec = D.batch_fields(nfields=16, lo=2 ** 12, hi=2 ** 22, seed=77)
    d_in = []
    for i, f in enumerate(spec):
        x = D.generate_gpu(f["kind"], f["dims"], f["seed"])
        if i % 4 == 1:
            x[7] = float("nan")
        if i % 4 == 2:
            x[x.numel() // 2] = 3e8
        d_in.append(x)
    ebs = [f["eb"] for f in spec]
    bp = fzgpu.BatchPlan([f["dims"] for f in spec], mode=fzgpu.MODE_OUTLIER)
    out, res = bp.compress_outlier(d_in, ebs)
    assert all(r["status"] == 0 for r in res)
    sides = out.side_channels()
    xs = bp.decompress_outlier([out.payloads[i][: res[i]["payload_bytes"]] for i in range(len(spec))],
                               [out.tails[i][: res[i]["tail_len"]] for i in range(len(spec))], sides,
                               [r["ebx2"] for r in res], [r["vmin"] for r in res])
    for i, x in enumerate(d_in):
        fin = torch.isfinite(x)
        xr = xs[i]
        assert torch.equal(xr[~fin].view(torch.int32), x[~fin].view(torch.int32)), i
        err = float((xr[fin].double() - x[fin].double()).abs().max())
        amax = float(x[fin].abs().max())
        assert err <= res[i]["eb_abs"] + 4 * 2.0 ** -24 * amax, (i, err, res[i]["eb_abs"])
