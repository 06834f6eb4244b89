"""Host-resident requests through the stager (pinned, asynchronous staging, fz_stager_*): the
same bytes as the device-resident entry points for every request in flight, in and out of
order, from pinned and pageable host memory; decompress staging; slot reuse and errors."""
import sys

import numpy as np
import pytest

import fzgpu
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(dims, k):
    rng = np.random.default_rng(sum(dims))
    n = dims[0] * dims[1] * dims[2]
    base = np.cumsum(rng.normal(size=n)).astype(np.float32)
    return [(base * np.float32(1 + 0.25 * i) + np.float32(i)).astype(np.float32) for i in range(k)]


@pytest.mark.parametrize("dims,depth", [((3600, 180, 1), 3), ((5000, 1, 1), 1), ((33, 20, 17), 2),
                                        ((64, 64, 64), 8)])
def test_staged_compress_equals_device_compress(dims, depth):
    import torch
    say(f"case stage compress dims={dims} depth={depth}")
    xs = _fields(dims, 7)
    bufs = []
    for x in xs:
        b = fzgpu.HostBuffer(x.size, np.float32)
        b.array[:] = x
        bufs.append(b)
    st = fzgpu.Stager(*dims, depth=depth)
    plan = fzgpu.Plan(*dims)
    tickets = []
    for i, b in enumerate(bufs):
        tickets.append(st.compress(b.array, 1e-3))
        # wait for the request that is about to lose its slot (its results stay readable until then)
        if len(tickets) >= depth:
            t = tickets[len(tickets) - depth]
            res, payload, tail = st.wait(t)
            ref_p, ref_t, ref_r = plan.compress(torch.from_numpy(xs[t]).to("cuda"), 1e-3)
            assert res["status"] == 0 and res["payload_bytes"] == ref_r["payload_bytes"]
            assert np.array_equal(payload, ref_p.cpu().numpy())
            assert np.array_equal(tail, ref_t.cpu().numpy().view(np.uint16))
    for t in tickets[max(0, len(tickets) - depth + 1):] if depth > 1 else []:
        res, payload, _ = st.wait(t)
        ref = O.compress(xs[t], dims, 1e-3)
        assert np.array_equal(payload, ref["payload"])
    # the oldest ticket's slot has been reused
    if len(tickets) > depth:
        with pytest.raises(fzgpu.FzError):
            st.wait(tickets[0])
    st.close()


def test_pageable_input_and_errors():
    dims = (2000, 50, 1)
    say("case stage pageable")
    x = _fields(dims, 1)[0]
    st = fzgpu.Stager(*dims, depth=2)
    t = st.compress(x, 1e-3)  # ordinary numpy memory: a blocking copy, same result
    res, payload, tail = st.wait(t)
    ref = O.compress(x, dims, 1e-3)
    assert np.array_equal(payload, ref["payload"]) and np.array_equal(tail, ref["tail"])
    c = np.full(x.size, 2.0, np.float32)
    res, _, _ = st.wait(st.compress(c, 1e-3))
    assert res["status"] == 4  # constant field
    with pytest.raises(fzgpu.FzError):
        st.wait(12345)
    st.close()


@pytest.mark.parametrize("dims", [(3600, 180, 1), (33, 20, 17), (5000, 1, 1)])
def test_staged_decompress(dims):
    say(f"case stage decompress dims={dims}")
    xs = _fields(dims, 4)
    refs = [O.compress(x, dims, 1e-3) for x in xs]
    st = fzgpu.Stager(*dims, depth=3)
    outs = [fzgpu.HostBuffer(xs[0].size, np.float32) for _ in xs]
    ins = []
    tickets = []
    for r, o in zip(refs, outs):
        p = fzgpu.HostBuffer(r["payload"].size, np.uint8)
        p.array[:] = r["payload"]
        tl = fzgpu.HostBuffer(max(r["tail"].size, 1), np.uint16)
        tl.array[: r["tail"].size] = r["tail"]
        ins.append((p, tl))
        tickets.append(st.decompress(p.array, tl.array[: r["tail"].size], r["info"]["ebx2"], o.array))
    for t, r, o in zip(tickets[-3:], refs[-3:], outs[-3:]):
        assert st.wait(t) is None
        rc, ref_out = O.decompress(r["payload"], r["tail"], xs[0].size, dims, r["info"]["ebx2"])
        assert np.array_equal(o.array.view(np.uint32), ref_out.view(np.uint32))
    st.close()
