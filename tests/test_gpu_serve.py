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
