"""fz_compress_f32_noalloc (no workspace argument, no allocation) against fz_compress_f32 through
the C ABI: payload, tail and result record equal byte for byte, on two calls in a row (the
second on a smaller field, which must not see anything the first left in the library's
array), plus the capacity refusal above FZ_NOALLOC_MAX_TILES."""
import ctypes
import sys

import numpy as np
import pytest

import fzgpu
from fzgpu import datagen as D

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(plan, d_in, eb):
    import torch
    L = fzgpu.lib()
    payload, tail, res = plan.alloc_outputs()
    payload.fill_(0xA5)  # stale bytes a wrong length would expose
    st = L.fz_compress_f32_noalloc(plan.handle, fzgpu._ptr(d_in), float(eb), fzgpu._ptr(payload), fzgpu._ptr(tail),
                                   fzgpu._ptr(res), fzgpu._stream())
    assert st == 0, st
    torch.cuda.synchronize()
    r = fzgpu.read_result(res)
    return payload[: r["payload_bytes"]], tail[: r["tail_len"]], r


def _same(plan, d_in, eb):
    import torch
    pa, ta, ra = plan.compress(d_in, eb)            # the workspace entry point
    pb, tb, rb = _noalloc(plan, d_in, eb)           # the no-allocation one
    assert ra == rb, (ra, rb)
    assert torch.equal(pa, pb) and torch.equal(ta, tb)
    return rb


@pytest.mark.parametrize("layout", [0, 1])
def test_noalloc_equals_the_workspace_entry_point_on_two_calls(layout):
    import torch
    assert fzgpu.lib().fz_noalloc_scratch_bytes() >= 4096 * 16000
    big_dims, small_dims = (3600, 1800, 1), (1000, 37, 1)
    x_big = torch.from_numpy(D.waves(big_dims, 11).ravel()).cuda()
    x_small = torch.from_numpy(D.peaks(small_dims, 12).ravel()).cuda()
    pb, ps = fzgpu.Plan(*big_dims), fzgpu.Plan(*small_dims)
    pb.set_tile_layout(layout)
    ps.set_tile_layout(layout)
    say(f"no-alloc compress {big_dims} then {small_dims}, layout {layout}")
    r1 = _same(pb, x_big, 1e-3)
    r2 = _same(ps, x_small, 1e-4)
    assert r1["nblocks"] > r2["nblocks"] > 0


def test_noalloc_3d_and_1d_then_smaller():
    import torch
    cases = [((500, 500, 100), "blobs", 1e-4), ((6_000_001, 1, 1), "waves", 1e-3), ((33, 20, 17), "blobs", 1e-3)]
    for dims, kind, eb in cases:
        say(f"no-alloc compress {dims} {kind} eb={eb}")
        x = D.generate_gpu(kind, dims, 21)
        _same(fzgpu.Plan(*dims), x, eb)
        del x
        torch.cuda.empty_cache()


def test_noalloc_refuses_fields_beyond_its_array():
    import torch
    dims = (2048 * 16001, 1, 1)
    plan = fzgpu.Plan(*dims, workspace=False)
    x = torch.empty(16, dtype=torch.float32, device="cuda")  # never read: refused before any launch
    payload = torch.empty(16, dtype=torch.uint8, device="cuda")
    L = fzgpu.lib()
    st = L.fz_compress_f32_noalloc(plan.handle, fzgpu._ptr(x), 1e-3, fzgpu._ptr(payload), fzgpu._ptr(payload),
                                   fzgpu._ptr(payload), fzgpu._stream())
    assert st == 7
    ok = fzgpu.Plan(2048 * 16000 + 2047, 1, 1, workspace=False)
    assert ok.sizes.tiles == 16000
