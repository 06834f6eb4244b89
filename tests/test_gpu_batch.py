"""Batched form: every field's bytes equal the oracle and the per-field GPU path.

The batch mixes 1D / 2D / 3D fields, fields shorter than one tile, fields with len > x*y*z,
and the per-field eb values {1e-2, 1e-3, 1e-4}.  Bit-exact bar for payload, tail codes,
result records and decompressed floats; the same bytes under every launch tuning and on
every replay of a captured graph.
"""
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


# (x, y, z) or (x, y, z, len)
MIX = [(5000, 1, 1), (4096, 1, 1), (2047, 1, 1), (9, 1, 1), (100, 77, 1), (16, 16, 1), (17, 33, 1), (3600, 20, 1),
       (1, 3000, 1), (33, 20, 17), (64, 64, 64), (500, 50, 10), (1, 1, 5000), (7, 1, 900), (70000, 1, 1),
       (300, 299, 1), (41, 37, 29), (1000, 3, 1, 6000), (2048, 1, 1, 4100), (3601, 37, 1), (9, 50, 40),
       (13, 1, 700), (200, 30, 12), (1000, 9, 9), (1100, 9, 9)]
EBS = (1e-2, 1e-3, 1e-4)


def _len(s):
    return s[3] if len(s) > 3 else s[0] * s[1] * s[2]


def _field(s, seed):
    rng = np.random.default_rng(seed)
    n = _len(s)
    x = (np.cumsum(rng.normal(size=n)) * (0.5 + seed % 3)).astype(np.float32)
    x[rng.random(n) < 0.3] = 0
    x[0] = 1.5
    return x


def _to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _np(t, dtype=None):
    a = t.cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def _batch(shapes, seed0=100):
    xs = [_field(s, seed0 + i) for i, s in enumerate(shapes)]
    ebs = [EBS[i % 3] for i in range(len(shapes))]
    return xs, ebs


def _field_bytes(out, res, i):
    r = res[i]
    return _np(out.payloads[i][: r["payload_bytes"]]), _np(out.tails[i][: r["tail_len"]], np.uint16)


@pytest.mark.parametrize("mode", [O.MODE_FIXED, O.MODE_LITERAL])
def test_batch_matches_oracle(mode):
    xs, ebs = _batch(MIX, 100 + mode)
    say(f"batch of {len(MIX)} fields, mode {mode}")
    plan = fzgpu.BatchPlan(MIX, mode=mode)
    out, res = plan.compress([_to_dev(x) for x in xs], ebs)
    for i, (s, x) in enumerate(zip(MIX, xs)):
        ref = O.compress(x, s[:3], ebs[i], mode)
        info = ref["info"]
        r = res[i]
        assert r["status"] == 0, (i, s)
        for k_gpu, k_ref in [("nblocks", "K"), ("payload_bytes", "payload_bytes"), ("overflow", "overflow"),
                             ("vmin", "vmin"), ("vmax", "vmax"), ("ebx2", "ebx2"), ("ebx2_r", "ebx2_r"),
                             ("eb_abs", "eb_abs")]:
            assert r[k_gpu] == info[k_ref], (i, s, k_gpu, r[k_gpu], info[k_ref])
        p, t = _field_bytes(out, res, i)
        assert np.array_equal(p, ref["payload"]), (i, s)
        if mode == O.MODE_FIXED:
            assert r["tail_len"] == info["tail_len"]
            assert np.array_equal(t, ref["tail"]), (i, s)
        else:
            assert r["tail_len"] == 0


def test_batch_equals_per_field_plain_and_fused():
    xs, ebs = _batch(MIX, 300)
    plan = fzgpu.BatchPlan(MIX)
    out, res = plan.compress([_to_dev(x) for x in xs], ebs)
    for i, (s, x) in enumerate(zip(MIX, xs)):
        say(f"per-field {s}")
        single = fzgpu.Plan(*s[:3], length=_len(s))
        for fused in (True, False):
            payload, tail, r = single.compress(_to_dev(x), ebs[i], fused=fused)
            assert r == res[i], (i, s, fused)
            p, t = _field_bytes(out, res, i)
            assert np.array_equal(p, _np(payload)) and np.array_equal(t, _np(tail, np.uint16)), (i, s, fused)


def test_same_bytes_under_every_tuning():
    xs, ebs = _batch(MIX, 500)
    plan = fzgpu.BatchPlan(MIX)
    d = [_to_dev(x) for x in xs]
    base_out, base_res = plan.compress(d, ebs)
    base = [_field_bytes(base_out, base_res, i) for i in range(len(MIX))]
    for chunk, xcd, gw, ng, wt in [(16384, (0, 0, 0), 1, 1, (1, 1, 1)), (262144, (1, 0, 1), 4, 2, (0, 1, 0)),
                                   (65536, (0, 1, 0), 2, 4, (1, 0, 1)), (16384, (1, 1, 1), 8, 16, (1, 1, 1)),
                                   (65536, (1, 1, 1), 16, 8, (0, 0, 0)), (65536, (1, 1, 1), 16, 1, (1, 1, 1))]:
        say(f"tuning chunk={chunk} xcd={xcd} gather_waves={gw} groups={ng} wave_tile={wt}")
        plan.set_tuning(range_chunk=chunk, xcd_order=xcd, gather_waves=gw, groups=ng, wave_tile=wt)
        out, res = plan.compress(d, ebs)
        assert res == base_res
        for i in range(len(MIX)):
            p, t = _field_bytes(out, res, i)
            assert np.array_equal(p, base[i][0]) and np.array_equal(t, base[i][1]), (i, chunk, xcd, gw, ng, wt)


def test_failed_fields_do_not_disturb_the_others():
    shapes = [(5000, 1, 1), (3000, 1, 1), (100, 77, 1), (33, 20, 17), (9000, 1, 1)]
    xs, ebs = _batch(shapes, 700)
    xs[1] = np.full(3000, 2.5, np.float32)  # constant field
    xs[3] = xs[3].copy()
    xs[3][777] = np.nan                     # non-finite value
    plan = fzgpu.BatchPlan(shapes)
    out, res = plan.compress([_to_dev(x) for x in xs], ebs)
    assert [r["status"] for r in res] == [0, 4, 0, 5, 0]
    for i in (0, 2, 4):
        ref = O.compress(xs[i], shapes[i], ebs[i], O.MODE_FIXED)
        p, t = _field_bytes(out, res, i)
        assert np.array_equal(p, ref["payload"]) and np.array_equal(t, ref["tail"]), i


def test_graph_replays_match_eager():
    import torch
    xs, ebs = _batch(MIX, 900)
    plan = fzgpu.BatchPlan(MIX)
    d = [_to_dev(x) for x in xs]
    eager_out, eager_res = plan.compress(d, ebs)
    eager = [_field_bytes(eager_out, eager_res, i) for i in range(len(MIX))]
    out = plan.alloc_outputs()
    plan.set_tuning(groups=4)  # the graph holds the three-stream schedule
    g = plan.compress_graph(d, ebs, out)
    single = fzgpu.Plan(100, 77, 1)
    for rep in range(3):
        say(f"graph replay {rep}")
        out.payload_arena.fill_(0xA5)
        out.res.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        res = fzgpu.read_results(out.res, len(MIX))
        assert res == eager_res, rep
        for i in range(len(MIX)):
            p, t = _field_bytes(out, res, i)
            assert np.array_equal(p, eager[i][0]) and np.array_equal(t, eager[i][1]), (rep, i)
        single.compress(d[4], 1e-3)  # unrelated eager work between replays
    g.close()


def test_batch_decompress_matches_oracle_and_graph():
    import torch
    xs, ebs = _batch(MIX, 1100)
    plan = fzgpu.BatchPlan(MIX)
    out, res = plan.compress([_to_dev(x) for x in xs], ebs)
    ebx2 = [r["ebx2"] for r in res]
    pays = [out.payloads[i] for i in range(len(MIX))]
    tails = [out.tails[i][: res[i]["tail_len"]] for i in range(len(MIX))]
    say("batched decompress")
    outs = plan.decompress(pays, tails, ebx2)
    for i, (s, x) in enumerate(zip(MIX, xs)):
        ref = O.compress(x, s[:3], ebs[i], O.MODE_FIXED)
        rc, xo = O.decompress(ref["payload"], ref["tail"], x.size, s[:3], ref["info"]["ebx2"])
        assert rc == 0
        n = s[0] * s[1] * s[2]
        xr = _np(outs[i])
        assert np.array_equal(xr.view(np.uint32), xo[:n].view(np.uint32)), (i, s)
        tol = res[i]["eb_abs"] + 4 * 2.0 ** -24 * float(np.abs(x[:n]).max())
        assert float(np.abs(xr - x[:n]).max()) <= tol, (i, s)
    say("batched decompress graph")
    gouts = [torch.empty(max(s[0] * s[1] * s[2], 4), dtype=torch.float32, device="cuda") for s in MIX]
    g = plan.decompress_graph(pays, tails, ebx2, gouts)
    for rep in range(2):
        for o in gouts:
            o.fill_(-7.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, s in enumerate(MIX):
            n = s[0] * s[1] * s[2]
            assert torch.equal(gouts[i][:n], outs[i]), (rep, i)
    g.close()


def test_autotune_keeps_the_bytes():
    xs, ebs = _batch(MIX, 1300)
    plan = fzgpu.BatchPlan(MIX)
    d = [_to_dev(x) for x in xs]
    out0, res0 = plan.compress(d, ebs)
    out = plan.alloc_outputs()
    t = plan.autotune(d, ebs, out, reps=2)
    say(f"autotune chose {t}")
    assert t["range_chunk"] in (16384, 65536, 262144) and t["gather_waves"] in (4, 8, 16)
    assert t["groups"] in (1, 2, 4, 8, 16)
    assert plan.tuning == t
    out1, res1 = plan.compress(d, ebs)
    assert res1 == res0
    for i in range(len(MIX)):
        assert all(np.array_equal(a, b) for a, b in zip(_field_bytes(out1, res1, i), _field_bytes(out0, res0, i)))


def test_uneven_batch_full_size_equals_per_field_plain():
    """The 256-field uneven batch (2.6 G floats, generated on the device) against the plain
    per-field pipeline, every byte; three small fields also against the oracle."""
    import torch
    spec = D.batch_fields()
    say("generating the 256-field batch on the device")
    d = [D.generate_gpu(f["kind"], f["dims"], f["seed"]) for f in spec]
    shapes = [f["dims"] for f in spec]
    ebs = [f["eb"] for f in spec]
    plan = fzgpu.BatchPlan(shapes)
    out, res = plan.compress(d, ebs)
    say("batched compress done; per-field plain pipeline")
    for i, f in enumerate(spec):
        single = fzgpu.Plan(*f["dims"])
        payload, tail, r = single.compress(d[i], ebs[i], fused=False)
        assert r == res[i], (i, f)
        n = r["payload_bytes"]
        assert torch.equal(out.payloads[i][:n], payload), (i, f)
        assert torch.equal(out.tails[i][: r["tail_len"]], tail), (i, f)
        del single, payload, tail
    # batched decompress of every field: the error bound, strict where the reference's binary32
    # arithmetic allows it (slack 4 * 2^-24 * max|x|, as for the single-field workloads)
    outs = plan.decompress([out.payloads[i][: res[i]["payload_bytes"]] for i in range(len(spec))],
                           [out.tails[i][: res[i]["tail_len"]] for i in range(len(spec))],
                           [res[i]["ebx2"] for i in range(len(spec))])
    strict = 0
    for i in range(len(spec)):
        n = shapes[i][0] * shapes[i][1] * shapes[i][2]
        err = float((outs[i][:n].double() - d[i].double()).abs().max())
        amax = float(d[i].abs().max())
        assert err <= res[i]["eb_abs"] + 4 * 2.0 ** -24 * amax, (i, spec[i])
        strict += err <= res[i]["eb_abs"]
    say(f"error bound: {strict} of {len(spec)} fields within eb (max - min) exactly")
    del outs
    small = sorted(range(len(spec)), key=lambda i: shapes[i][0] * shapes[i][1] * shapes[i][2])[:3]
    for i in small:
        say(f"oracle check of field {i} {shapes[i]}")
        x = _np(d[i])
        ref = O.compress(x, shapes[i], ebs[i], O.MODE_FIXED)
        p, t = _field_bytes(out, res, i)
        assert np.array_equal(p, ref["payload"]) and np.array_equal(t, ref["tail"]), i


def test_batch_with_the_largest_field():
    """A batch holding the largest field a plan accepts (17.2 GB, over 2^20 units) between two
    small ones: every field's archive equals the single-field fused path's, and the batched
    decompress equals the single-field decompress bit for bit."""
    import torch
    big = (0xFFFFF000 - 100, 1, 1)
    shapes = [(1000, 999, 1), big, (33, 20, 17)]
    torch.cuda.empty_cache()
    say(f"batch with the largest field {big}")
    xs = [_to_dev(_field(shapes[0], 41)), torch.arange(big[0], dtype=torch.float32, device="cuda"),
          _to_dev(_field(shapes[2], 43))]
    xs[1].mul_(1e-3).sin_()
    ebs = [1e-3, 1e-4, 1e-2]
    plan = fzgpu.BatchPlan(shapes)
    out, res = plan.compress(xs, ebs)
    singles = []
    for i, s in enumerate(shapes):
        single = fzgpu.Plan(*s)
        payload, tail, r = single.compress(xs[i], ebs[i])
        assert r == res[i], (i, s)
        assert torch.equal(out.payloads[i][: r["payload_bytes"]], payload), (i, s)
        assert torch.equal(out.tails[i][: r["tail_len"]], tail), (i, s)
        singles.append((single, payload, tail, r))
    pays = [out.payloads[i] for i in range(3)]
    tails = [out.tails[i][: res[i]["tail_len"]] for i in range(3)]
    outs = plan.decompress(pays, tails, [r["ebx2"] for r in res])
    for i, (single, payload, tail, r) in enumerate(singles):
        one = single.decompress(payload, tail, r["ebx2"])
        assert torch.equal(outs[i].view(torch.int32), one.view(torch.int32)), i
        del one
    del xs, out, outs, singles, plan
    torch.cuda.empty_cache()


def _random_batch(seed, nfields):
    """Seeded random shapes for a batch: a third each 1D / 2D / 3D, 2 .. 600K floats, any
    residues; the 3D rows span every fused-window width (X from 1 to 1500)."""
    rng = np.random.default_rng(4000 + seed)
    shapes = []
    for i in range(nfields):
        nd = 1 + i % 3
        total = int(np.exp(rng.uniform(0, np.log(6e5))))
        if nd == 1:
            shapes.append((max(total, 2), 1, 1))
        elif nd == 2:
            x = int(np.exp(rng.uniform(0, np.log(max(total, 2)))))
            shapes.append((max(x, 1), max(2, total // max(x, 1)), 1))
        else:
            x = int(np.exp(rng.uniform(0, np.log(1500))))
            y = int(np.exp(rng.uniform(0, np.log(max(total // x, 2)))))
            shapes.append((x, max(y, 1), max(2, total // max(x * y, 1))))
    return shapes


@pytest.mark.parametrize("seed", range(4))
def test_random_batches_equal_single_field(seed):
    """Seeded random batches of 48 fields: every field's archive equals the single-field fused
    compress, and the batched decompress (3D fields in fused windows of every size class)
    equals the single-field decompress, bit for bit."""
    shapes = _random_batch(seed, 48)
    xs = [_field(s, 9000 + 64 * seed + i) for i, s in enumerate(shapes)]
    ebs = [float(10.0 ** np.random.default_rng(seed * 100 + i).uniform(-4.5, -1.5)) for i in range(len(shapes))]
    say(f"random batch seed={seed}: {len(shapes)} fields, 3D X {sorted(s[0] for s in shapes if s[2] > 1)}")
    plan = fzgpu.BatchPlan(shapes)
    out, res = plan.compress([_to_dev(x) for x in xs], ebs)
    assert all(r["status"] == 0 for r in res)
    pays = [out.payloads[i] for i in range(len(shapes))]
    tails = [out.tails[i][: res[i]["tail_len"]] for i in range(len(shapes))]
    outs = plan.decompress(pays, tails, [r["ebx2"] for r in res])
    for i, (s, x) in enumerate(zip(shapes, xs)):
        single = fzgpu.Plan(*s)
        payload, tail, r = single.compress(_to_dev(x), ebs[i])
        assert r == res[i], (i, s)
        p, t = _field_bytes(out, res, i)
        assert np.array_equal(p, _np(payload)) and np.array_equal(t, _np(tail, np.uint16)), (i, s)
        one = single.decompress(payload, tail, r["ebx2"])
        assert np.array_equal(_np(one).view(np.uint32), _np(outs[i]).view(np.uint32)), (i, s)
