"""Serving loop on the device: requests that share batches and overlap on the GPU give the
same bytes as each request served alone (compress: payload, tail and result record;
decompress: every float bit-exact, and within the error bound of the original field)."""
import sys

import numpy as np
import pytest

import fzgpu
from fzgpu import datagen as D
from fzgpu import serve as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pool():
    shapes = [(65536, 1, 1), (300, 301, 1), (41, 37, 29), (1 << 20, 1, 1), (1024, 700, 1), (64, 64, 64),
              (123457, 1, 1), (2001, 333, 1), (97, 55, 33), (500000, 1, 1), (4096, 64, 1), (100, 100, 50)]
    return [(d, D.generate_gpu(D.BATCH_KINDS[i % 3], d, 50 + i)) for i, d in enumerate(shapes)]


def test_served_results_equal_alone():
    import torch
    pool = _pool()
    reqs = S.make_requests(3000.0, 0.03, pool, seed=11)
    say(f"serving {len(reqs)} requests")
    srv = S.Server(S.DeviceBackend(nstreams=3), slots=3, max_fields=4, max_archives=10_000)
    done = srv.run(reqs, realtime=True, on_arrival=S.pick_archive(12))
    assert len(done) == len(reqs)
    say(f"batches {len(srv.batches)}, max in flight {srv.max_inflight}, "
        f"decompress {sum(r.kind == 'decompress' for r in done)}")
    src_of = {}
    for r in done:
        if r.kind == "compress":
            src_of[id(r.result)] = r
    for r in done:
        plan = fzgpu.Plan(*r.dims)
        if r.kind == "compress":
            payload, tail, rec = plan.compress(r.source, r.eb)
            a = r.result
            assert a.record == rec, r.rid
            assert torch.equal(a.payload, payload) and torch.equal(a.tail, tail), r.rid
        else:
            a = r.source
            alone = plan.decompress(a.payload, a.tail, a.record["ebx2"])
            assert torch.equal(r.result, alone), r.rid
            x = src_of[id(a)].source.float()
            tol = a.record["eb_abs"] + 4 * 2.0 ** -24 * float(x.abs().max())
            assert float((r.result - x).abs().max()) <= tol, r.rid


def _copy(a):
    """An Archive's tensors copied out of the server's memory (which it may reuse once it has
    dropped the archive)."""
    return S.Archive(a.dims, a.payload.clone(), a.tail.clone(), dict(a.record),
                     side=a.side.clone() if a.side is not None else None)


def _check_alone(done, kept, orig):
    """kept: rid -> copy of the Archive a compress request produced, made when it finished;
    orig: id of that served Archive -> rid (a decompress request's source is a served Archive)."""
    import torch
    src_of = {rid: r for r in done for rid in [r.rid] if r.kind == "compress"}
    for r in done:
        plan = fzgpu.Plan(*r.dims)
        if r.kind == "compress":
            payload, tail, rec = plan.compress(r.source, r.eb)
            a = kept[r.rid]
            assert a.record == rec, r.rid
            assert torch.equal(a.payload, payload) and torch.equal(a.tail, tail), r.rid
        else:
            rid = orig[id(r.source)]
            a = kept[rid]
            alone = plan.decompress(a.payload, a.tail, a.record["ebx2"])
            assert torch.equal(r.result, alone), r.rid
            x = src_of[rid].source.float()
            tol = a.record["eb_abs"] + 4 * 2.0 ** -24 * float(x.abs().max())
            assert float((r.result - x).abs().max()) <= tol, r.rid


@pytest.mark.parametrize("arena", [0, 3_000_000_000])
def test_saturating_stream_shares_batches_under_the_budget(arena):
    """Arrivals far faster than the device serves them (one every 50 us), over a pool whose
    largest field is 268 MB (every 8th request), and a 0.6 GB budget smaller than what the waiting requests need
    together: waiting requests must share batches, several batches must be on the device at
    once, and admission must stop at the budget.  Every request's result still equals the same
    request served alone, byte for byte."""
    big = (8192, 8192, 1)  # 67.1M floats, 268 MB
    shapes = [big, (1 << 20, 1, 1), (300, 301, 1), (64, 64, 64), (2001, 333, 1), (123457, 1, 1),
              (97, 55, 33), (4096, 64, 1)]
    pool = [(d, D.generate_gpu(D.BATCH_KINDS[i % 3], d, 70 + i)) for i, d in enumerate(shapes)]
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(240):
        dims, x = pool[0] if i % 8 == 0 else pool[1 + int(rng.integers(0, len(pool) - 1))]
        kind = "decompress" if i % 5 == 4 else "compress"
        reqs.append(S.Request(i, kind, dims, S.SERVE_EBS[i % 3], x, 5e-5 * i))
    budget = 0.6e9
    need_big = sum(S.request_bytes(S.Request(0, "compress", big)))
    assert 2 * need_big < budget < sum(sum(S.request_bytes(r)) for r in reqs)
    say(f"saturating stream: {len(reqs)} requests, budget {budget / 1e9} GB, big request {need_big / 1e6:.0f} MB")
    be = S.DeviceBackend(nstreams=3, arena_bytes=arena)
    srv = S.Server(be, hbm_budget_bytes=budget, slots=3, max_fields=16, max_archives=24)
    # archives to decompress from the first arrival on: every pool field served once beforehand
    kept, orig = {}, {}

    def keep(r):  # the archive as served, copied before any eviction can drop it
        if r.kind == "compress":
            kept[r.rid] = _copy(r.result)
            orig[id(r.result)] = r.rid

    warm = [S.Request(1000 + i, "compress", d, 1e-3, x, 0.0) for i, (d, x) in enumerate(pool)]
    srv.run(warm, realtime=False, on_complete=keep)
    srv.completed.clear()
    srv.batches.clear()
    srv.max_inflight = srv.budget_limited = 0
    done = srv.run(reqs, realtime=True, on_arrival=S.pick_archive(3), on_complete=keep)
    assert len(done) == len(reqs)
    sizes = [b[2] for b in srv.batches]
    st = S.summarize(srv, done)
    say(f"batches {len(sizes)}, max batch {max(sizes)}, mean {st['mean_batch']}, max in flight {srv.max_inflight}, "
        f"budget-limited {srv.budget_limited}, peak {srv.budget.peak / 1e9:.3f} GB, "
        f"decompress {st['decompress']}, hist {st['batch_size_hist']}")
    assert max(sizes) >= 2                      # requests shared a batch
    assert srv.max_inflight >= 2                # batches overlapped on the device
    assert srv.budget_limited >= 1              # admission stopped at the budget
    assert srv.budget.peak <= srv.budget.total
    assert sum(r.dims == big for r in done) >= 5
    assert st["decompress"] >= 1
    if arena:  # workspaces and archives came from the arena, and its blocks came back
        say(f"arena peak {be.arena.peak / 1e9:.3f} GB, misses {be.arena.misses}")
        assert be.arena.peak > 0 and be.arena.misses == 0
        srv.close()  # every archive dropped: only the pooled plans' workspaces stay in the arena
        pooled = sum(b[1] for lst in be.pool.free.values() for p in lst for b in p.arena_blocks)
        assert be.arena.used == pooled
    _check_alone(warm + done, kept, orig)


def test_outlier_mode_serving_handles_fields_the_fixed_mode_refuses():
    """A served stream in the outlier mode over a pool with a constant field, a field with NaN
    and +-Inf and a field with deltas beyond the u16 code: every request completes (status 0)
    and equals the single-field outlier entry point on the same input; decompress requests
    restore the non-finite values bit for bit."""
    import torch
    pool = []
    rng = np.random.default_rng(9)
    for i, d in enumerate([(300, 301, 1), (65536, 1, 1), (41, 37, 29), (2001, 333, 1)]):
        x = D.generate_gpu(D.BATCH_KINDS[i % 3], d, 90 + i)
        if i == 0:
            x.fill_(1.75)
        elif i == 1:
            x[5] = float("nan")
            x[9] = float("inf")
            x[11] = float("-inf")
        elif i == 2:
            x[rng.integers(0, x.numel(), 4)] = 1e7
        pool.append((d, x))
    reqs = S.make_requests(4000.0, 0.02, pool, seed=21)
    srv = S.Server(S.DeviceBackend(mode=fzgpu.MODE_OUTLIER, nstreams=3), slots=3, max_fields=8, max_archives=10_000)
    kept = {}
    done = srv.run(reqs, realtime=True, on_arrival=S.pick_archive(22),
                   on_complete=lambda r: kept.__setitem__(r.rid, r.result) if r.kind == "compress" else None)
    assert len(done) == len(reqs)
    say(f"outlier serving: {len(done)} requests, {len(srv.batches)} batches, max in flight {srv.max_inflight}")
    src_of = {id(kept[r.rid]): r for r in done if r.kind == "compress"}
    for r in done:
        plan = fzgpu.Plan(*r.dims, mode=fzgpu.MODE_OUTLIER)
        if r.kind == "compress":
            a = kept[r.rid]
            assert a.record["status"] == 0, r.rid
            p, t, rec, side = plan.compress_outlier(r.source, r.eb)
            assert a.record == rec, r.rid
            assert torch.equal(a.payload, p) and torch.equal(a.tail, t), r.rid
            assert torch.equal(a.side[: side.numel()], side), r.rid
        else:
            a = r.source
            x = src_of[id(a)].source
            fin = torch.isfinite(x)
            assert torch.equal(r.result[~fin].view(torch.int32), x[~fin].view(torch.int32)), r.rid
            if a.record["ebx2"] == 0.0:
                assert torch.equal(r.result, x), r.rid
            else:
                err = float((r.result[fin].double() - x[fin].double()).abs().max())
                assert err <= a.record["eb_abs"] + 4 * 2.0 ** -24 * float(x[fin].abs().max()), r.rid
