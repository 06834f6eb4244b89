"""Scheduler rules of the serving loop, on the CPU with a simulated backend: FIFO admission
per kind, the HBM budget and the slot limit are never exceeded, archives are evicted to make
room, latencies are consistent, and an oversized request is refused."""
import math

import numpy as np
import pytest

from fzgpu import serve as S


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-6  # a spinning scheduler loop takes time too
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeBackend:
    """A batch takes 1 ms + 1 ns per input byte of simulated time; the clock advances on poll."""

    def __init__(self, clock, server_ref):
        self.clock = clock
        self.server_ref = server_ref
        self.launched = []

    def launch(self, kind, reqs):
        srv = self.server_ref[0]
        assert srv.budget.used <= srv.budget.total
        self.launched.append((kind, [r.rid for r in reqs]))
        return {"end": self.clock.t + 1e-3 + 1e-9 * sum(r.input_bytes for r in reqs), "kind": kind}

    def poll(self, h):
        self.clock.t += 1e-5
        return self.clock.t >= h["end"]

    def finish(self, h, reqs):
        for r in reqs:
            r.result = S.Archive(r.dims, None, None, {"payload_bytes": 10, "tail_len": 0}) if h["kind"] == "compress" \
                else "floats"


def _server(budget=200e9, slots=3, max_fields=256, max_archives=64):
    clock = FakeClock()
    ref = [None]
    be = FakeBackend(clock, ref)
    srv = S.Server(be, hbm_budget_bytes=budget, slots=slots, max_fields=max_fields, max_archives=max_archives,
                   clock=clock, sleep=clock.sleep)
    ref[0] = srv
    return srv, be, clock


def _reqs(n, seed=0, dims=(1 << 20, 1, 1), decompress=False):
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(n):
        t += rng.exponential(1e-4)
        out.append(S.Request(i, "decompress" if decompress and i % 5 == 4 else "compress", dims, 1e-3, None, t))
    return out


def test_fifo_order_and_all_complete():
    srv, be, _ = _server(slots=2, max_fields=4)
    reqs = _reqs(40)
    done = srv.run(reqs, realtime=True)
    assert len(done) == 40
    order = [rid for kind, ids in be.launched for rid in ids]
    assert order == sorted(order)
    assert srv.max_inflight <= 2
    assert all(len(ids) <= 4 for _, ids in be.launched)
    for r in done:
        assert r.arrival <= r.admitted <= r.done


def test_budget_limits_admission():
    one = sum(S.request_bytes(S.Request(0, "compress", (1 << 20, 1, 1))))
    srv, be, _ = _server(budget=3.5 * one, slots=8, max_archives=0)
    srv.run(_reqs(20), realtime=False)
    assert all(len(ids) <= 3 for _, ids in be.launched)
    assert srv.budget.peak <= srv.budget.total
    assert srv.budget.used == 0


def test_archives_are_evicted_to_make_room():
    one = sum(S.request_bytes(S.Request(0, "compress", (1 << 20, 1, 1))))
    srv, _, _ = _server(budget=4 * one, slots=1, max_fields=2, max_archives=100)
    srv.run(_reqs(30), realtime=False)
    assert len(srv.completed) == 30
    assert srv.budget.peak <= srv.budget.total
    assert len(srv.archives) >= 1
    held = {id(b): b.nbytes for _, b in srv.archives}
    assert sum(held.values()) == srv.budget.used <= srv.budget.total


def test_archive_bytes_return_with_the_last_archive_of_a_batch():
    """The archives of one compress batch share one output buffer: evicting some of them frees
    nothing, evicting the last gives the whole buffer back."""
    out = S.request_bytes(S.Request(0, "compress", (1 << 20, 1, 1)))[1]
    srv, be, _ = _server(slots=1, max_fields=4, max_archives=100)
    reqs = [S.Request(i, "compress", (1 << 20, 1, 1), 1e-3, None, 0.0) for i in range(4)]
    srv.run(reqs, realtime=False)
    assert len(be.launched) == 1 and len(srv.archives) == 4
    assert srv.budget.used == 4 * out
    for k in range(3):
        assert srv._evict()
        assert srv.budget.used == 4 * out, k
    assert srv._evict() and srv.budget.used == 0 and not srv.archives
    assert not srv._evict()


def test_archives_read_by_waiting_decompress_requests_are_not_evicted():
    one = sum(S.request_bytes(S.Request(0, "compress", (1 << 20, 1, 1))))
    srv, be, _ = _server(budget=3 * one, slots=1, max_fields=1, max_archives=100)
    srv.run([S.Request(0, "compress", (1 << 20, 1, 1), 1e-3, None, 0.0)], realtime=False)
    a = srv.archives[0][0].result
    dec = S.Request(1, "decompress", (1 << 20, 1, 1), 1e-3, a, 0.0)
    srv.submit(dec)
    assert a.pins == 1 and not srv._evict()  # the only archive is pinned
    # compress requests that need the room must wait for the decompress instead of evicting
    more = [S.Request(2 + i, "compress", (1 << 20, 1, 1), 1e-3, None, 0.0) for i in range(3)]
    srv.run(more, realtime=False)
    assert dec.result == "floats" and a.pins == 0
    assert srv.budget.peak <= srv.budget.total


def test_decompress_requests_pick_archives():
    srv, be, _ = _server(slots=3)
    reqs = _reqs(50, seed=3, decompress=True)
    srv.run(reqs, realtime=False, on_arrival=S.pick_archive(5))
    kinds = {k for k, _ in be.launched}
    assert "compress" in kinds
    assert all(r.kind == "compress" or isinstance(r.source, S.Archive) for r in reqs)


def test_oversized_request_is_refused():
    srv, _, _ = _server(budget=1e6)
    with pytest.raises(ValueError):
        srv.run([S.Request(0, "compress", (1 << 22, 1, 1))], realtime=False)


def test_latency_stats_and_poisson():
    rng = np.random.default_rng(1)
    t = S.poisson_arrivals(1000.0, 2.0, rng)
    assert 1800 < len(t) < 2200 and all(a < b for a, b in zip(t, t[1:]))
    reqs = [S.Request(i, "compress", (10, 1, 1), arrival=0.0, done=(i + 1) * 1e-3) for i in range(100)]
    st = S.latency_stats(reqs)
    assert st["n"] == 100 and math.isclose(st["p50_ms"], 50.5, rel_tol=1e-6) and st["p99_ms"] <= 100.0


def test_plan_pool_reuses_by_shape_and_drops_least_recent():
    closed = []
    pool = S.PlanPool(100, nbytes=lambda p: p[1], close=closed.append)
    a, b, c, big = ("a", 40), ("b", 40), ("c", 40), ("big", 101)
    assert pool.take(((8, 1, 1),)) is None
    pool.give(((8, 1, 1),), a)
    pool.give(((9, 1, 1),), b)
    assert pool.bytes == 80
    assert pool.take(((8, 1, 1),)) is a and pool.bytes == 40  # handed out: no longer held
    pool.give(((8, 1, 1),), a)                                 # (9,) is now the least recent
    pool.give(((7, 1, 1),), c)                                 # 120 > 100: drops b
    assert closed == [b] and pool.bytes == 80
    assert pool.take(((9, 1, 1),)) is None
    pool.give(((6, 1, 1),), big)                               # larger than the whole pool
    assert closed == [b, big] and pool.bytes == 80
    assert pool.take(((7, 1, 1),)) is c and pool.take(((8, 1, 1),)) is a and pool.bytes == 0


def test_pinned_archives_never_block_admission_for_good():
    """A saturating stream over a pool with a field near half the budget, decompress requests
    reading archives that are pinned until they run: the archives offered to new decompress
    requests leave room for the largest request, so every request is served (a stall would
    raise) and the budget holds."""
    big = (8192, 8192, 1)
    shapes = [big, (1 << 20, 1, 1), (300, 301, 1), (64, 64, 64), (2001, 333, 1), (123457, 1, 1)]
    for seed in range(2):
        rng = np.random.default_rng(seed)
        reqs = []
        for i in range(240):
            dims = big if i % 8 == 0 else shapes[1 + int(rng.integers(0, len(shapes) - 1))]
            reqs.append(S.Request(i, "decompress" if i % 5 == 4 else "compress", dims, 1e-3, None, 5e-5 * i))
        srv, be, clock = _server(budget=0.6e9, slots=3, max_fields=16, max_archives=24)
        srv.run([S.Request(1000 + i, "compress", d, 1e-3, None, 0.0) for i, d in enumerate(shapes)], realtime=False)
        srv.completed.clear()
        done = srv.run(reqs, realtime=True, on_arrival=S.pick_archive(3))
        assert len(done) == len(reqs)
        assert sum(r.kind == "decompress" for r in done) > 0
        assert srv.budget.peak <= srv.budget.total


def test_a_request_that_can_never_fit_raises_instead_of_spinning():
    one = sum(S.request_bytes(S.Request(0, "compress", (1 << 20, 1, 1))))
    srv, _, _ = _server(budget=1.5 * one, slots=1, max_fields=1, max_archives=100)
    srv.run([S.Request(0, "compress", (1 << 20, 1, 1), 1e-3, None, 0.0)], realtime=False)
    a = srv.archives[0][0].result
    srv.submit(S.Request(1, "decompress", (1 << 20, 1, 1), 1e-3, a, 0.0))  # pins the archive
    srv.budget.take(srv.budget.total - srv.budget.used)  # memory held outside the server's requests
    with pytest.raises(RuntimeError):
        srv.run([], realtime=False)


# ---------------------------------------------------------------- device arena

def test_arena_best_fit_merges_and_counts():
    a = S.DeviceArena(1 << 20, device=None)
    b1 = a.alloc(1000)            # rounded to 1024
    b2 = a.alloc(256)
    b3 = a.alloc(4096)
    assert b1 == (0, 1024) and b2 == (1024, 256) and b3 == (1280, 4096)
    assert a.used == 1024 + 256 + 4096
    a.release(b2)                 # a 256-byte hole between b1 and b3
    assert a.alloc(200) == (1024, 256)  # best fit takes the hole, not the big tail block
    a.release((1024, 256))
    a.release(b1)                 # merges with the hole
    assert a.free[0] == (0, 1280)
    a.release(b3)                 # merges everything back into one block
    assert a.free == [(0, 1 << 20)] and a.used == 0
    assert a.peak == 1024 + 256 + 4096
    assert a.alloc(2 << 20) is None
    with pytest.raises(ValueError):
        a.release((0, 256))       # not allocated


def test_arena_random_alloc_free_keeps_its_invariants():
    rng = np.random.default_rng(5)
    a = S.DeviceArena(64 << 20, device=None)
    live = []
    for _ in range(3000):
        if live and (rng.random() < 0.45 or len(live) > 60):
            a.release(live.pop(int(rng.integers(0, len(live)))))
        else:
            b = a.alloc(int(rng.integers(1, 4 << 20)))
            if b is not None:
                live.append(b)
        # free blocks sorted, disjoint, never adjacent; used + free = size; live blocks disjoint
        offs = sorted(live + a.free)
        assert all(o1 + n1 <= o2 for (o1, n1), (o2, _) in zip(offs, offs[1:]))
        assert all(o1 + n1 < o2 for (o1, n1), (o2, _) in zip(a.free, a.free[1:]))
        assert a.used == sum(n for _, n in live) and a.used + sum(n for _, n in a.free) == a.size
    for b in live:
        a.release(b)
    assert a.free == [(0, a.size)]


def test_released_outputs_go_back_with_the_last_archive_of_their_batch():
    srv, be, _ = _server(max_fields=4, max_archives=6)
    freed = []
    be.free_outputs = freed.append       # the device backend returns the batch's arena blocks

    def finish(h, reqs, old=be.finish):
        old(h, reqs)
        keep = object()
        for r in reqs:
            if isinstance(r.result, S.Archive):
                r.result.keep = keep
    be.finish = finish
    reqs = _reqs(24)
    srv.run(reqs, realtime=False)
    kept = {id(r.result.keep) for r, _ in srv.archives}
    # every batch whose archives were all evicted has been released once, the others not yet
    assert len(freed) == len({id(k) for k in freed})
    assert not kept & {id(k) for k in freed}
    assert len(srv.archives) <= 6
    srv.close()
    assert not srv.archives and {id(k) for k in freed} >= kept
