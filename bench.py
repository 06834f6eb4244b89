#!/usr/bin/env python3
"""FZ compress benchmark on one MI355X.

A step is one compress of one field through the fused HIP path (device value range ->
fused dual-quant / bitshuffle / flags / look-back scan / encode), with the input already
resident in HBM.  The default workload is the 3600x1800 CESM-like 2D field at eb = 1e-3
(value-range relative).  Inputs rotate over enough copies (>= 1.2 GB) that each step reads
its field from HBM, not from the 256 MB MALL.

Printed: one JSON line with the headline (input GB/s), the compression ratio, the live
per-kernel roofline of the dominant kernel (HIP events on the launch stream) and the CPU
baseline (the OpenMP oracle on the host cores, bounded sample).  Extra workloads and the
plain per-stage pipeline are reported under "extra".
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fz-gpu_amd"))
# thread placement of the CPU baseline's OpenMP runtime, fixed before any library starts one
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# the process's CPU set before an OpenMP runtime binds this thread to one core
AFFINITY_AT_START = len(os.sched_getaffinity(0))

HBM_PEAK_GBPS = 8000.0
WORKLOAD_TEXT = {
    "cesm-like-3600x1800": "CESM-like 2D field: 32 plane waves, 1/k amplitudes, noise, 15% exact zeros in blobs",
    "detector-39200x388": "detector-like field, 97% exact zeros, Gaussian peaks + noisy rows",
    "blobs3d-500x500x100": "3D Gaussian blobs, 70% exact zeros",
}
METRIC = "compress GB/s + ratio at eb=1e-3 rel; batched vs per-field; p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(fn, steps):
    import torch
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for k in range(steps):
        fn(k)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def kernel_profile(plan, fn, reps):
    """Average per-kernel ms over `reps` calls (events recorded between the kernels)."""
    plan.profile(True)
    acc = None
    for k in range(reps):
        fn(k)
        t = plan.kernel_times()
        acc = t if acc is None else [a + b for a, b in zip(acc, t)]
    plan.profile(False)
    return [a / reps for a in acc]


def bench_field(x, dims, eb, steps, warmup, fused=True, rotate_bytes=1.2e9, label="", wave_tile=0):
    import numpy as np
    import torch

    import fzgpu
    n = x.size
    nbytes = 4 * n
    ncopies = max(2, min(64, math.ceil(rotate_bytes / nbytes)))
    base = torch.from_numpy(np.ascontiguousarray(x.ravel())).to("cuda")
    copies = [base] + [base.clone() for _ in range(ncopies - 1)]
    plan = fzgpu.Plan(*dims)
    plan.set_tile_layout(wave_tile)
    payload, tail, res = plan.alloc_outputs()

    def step(k):
        plan.compress_async(copies[k % ncopies], eb, payload, tail, res, fused=fused)

    for k in range(warmup):
        step(k)
    r = fzgpu.read_result(res)
    if r["status"] != 0:
        raise RuntimeError(f"{label}: compress failed with status {r['status']}")
    ms, wall = time_steps(step, steps)
    warm_ms, _ = time_steps(lambda k: plan.compress_async(base, eb, payload, tail, res, fused=fused), steps)
    kt = kernel_profile(plan, step, min(steps, 50))
    out = dict(n=n, ms=ms, wall_s=wall, warm_ms=warm_ms, kernel_ms=kt, payload_bytes=r["payload_bytes"],
               K=r["nblocks"], cr=nbytes / r["payload_bytes"], gbps=nbytes / (ms * 1e-3) / 1e9,
               warm_gbps=nbytes / (warm_ms * 1e-3) / 1e9, copies=ncopies)
    # decompress of the same archive
    p, t, r2 = plan.compress(base, eb, fused=True)
    dst = torch.empty(n, dtype=torch.float32, device="cuda")
    dms, _ = time_steps(lambda k: plan.decompress(p, t, r2["ebx2"], out=dst), steps)
    out["decompress_ms"] = dms
    out["decompress_gbps"] = nbytes / (dms * 1e-3) / 1e9
    out["decompress_kernel_ms"] = kernel_profile(plan, lambda k: plan.decompress(p, t, r2["ebx2"], out=dst),
                                                 min(steps, 50))
    # decompress + quality metrics against the original (one call, one stream)
    pa = p.clone() if p.data_ptr() % 16 else p
    d_stats = torch.zeros(160, dtype=torch.uint8, device="cuda")
    vfn = lambda k: plan.decompress_verify_async(pa, t, r2["ebx2"], base, dst, d_stats)
    vms, _ = time_steps(vfn, steps)
    out["decompress_verify_ms"] = vms
    out["decompress_verify_kernel_ms"] = kernel_profile(plan, vfn, min(steps, 50))
    out["verify_stats"] = fzgpu.read_stats(d_stats)
    del copies, base
    torch.cuda.empty_cache()
    return out


def _graph_of(fn):
    """torch CUDA graph of fn() (a loop of per-field library calls on the current stream)."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    torch.cuda.synchronize()
    return g


def bench_batch(steps, warmup, nfields=256, host_out=None):
    """The uneven batch: nfields fields (1D/2D/3D, 64K-64M floats, eb in {1e-2,1e-3,1e-4}),
    generated on the device.  Batched = one captured graph of the batched kernels; baseline =
    the plain per-field pipeline, launched field by field from the host and also captured as
    one graph (no host launch cost).  Every field's bytes are compared."""
    import torch

    import fzgpu
    from fzgpu import datagen as D
    spec = D.batch_fields(nfields)
    shapes = [f["dims"] for f in spec]
    ebs = [f["eb"] for f in spec]
    t0 = time.perf_counter()
    d = [D.generate_gpu(f["kind"], f["dims"], f["seed"]) for f in spec]
    torch.cuda.synchronize()
    log(f"  generated {nfields} fields in {time.perf_counter() - t0:.1f} s")
    in_bytes = sum(4 * t.numel() for t in d)
    plan = fzgpu.BatchPlan(shapes)
    out = plan.alloc_outputs()
    tune = plan.autotune(d, ebs, out, reps=3)
    # group-count sweep: eager launches and graph replays
    sweep = {}
    for ng in (1, 2, 4, 8, 16):
        plan.set_tuning(groups=ng)
        plan.compress_async(d, ebs, out)
        ms_e, _ = time_steps(lambda k: plan.compress_async(d, ebs, out), max(3, steps // 2))
        gg = plan.compress_graph(d, ebs, out)
        gg.launch()
        ms_g, _ = time_steps(lambda k: gg.launch(), max(3, steps // 2))
        gg.close()
        sweep[str(ng)] = {"eager_ms": round(ms_e, 4), "graph_ms": round(ms_g, 4)}
    plan.set_tuning(groups=tune["groups"])
    g = plan.compress_graph(d, ebs, out)
    for _ in range(warmup):
        g.launch()
    ms_batch, _ = time_steps(lambda k: g.launch(), steps)
    plan.profile(True)
    kt = [0.0] * 7
    for _ in range(3):
        plan.compress_async(d, ebs, out)
        kt = [a + b / 3 for a, b in zip(kt, plan.kernel_times())]
    plan.profile(False)
    res = fzgpu.read_results(out.res, nfields)
    assert all(r["status"] == 0 for r in res)
    pay_bytes = sum(r["payload_bytes"] for r in res)
    # per-field plain pipeline (and the fused single-field path) on the same inputs
    plans = [fzgpu.Plan(*s) for s in shapes]
    outs = [p.alloc_outputs() for p in plans]

    def per_field(fused):
        for i in range(nfields):
            plans[i].compress_async(d[i], ebs[i], *outs[i], fused=fused)

    pf_steps = max(2, steps // 4)
    per_field(False)
    ms_plain_host, _ = time_steps(lambda k: per_field(False), pf_steps)
    mismatch = 0
    for i in range(nfields):
        r = fzgpu.read_result(outs[i][2])
        n = r["payload_bytes"]
        same = (r == res[i] and torch.equal(outs[i][0][:n], out.payloads[i][:n]) and
                torch.equal(outs[i][1][: r["tail_len"]], out.tails[i][: r["tail_len"]]))
        mismatch += 0 if same else 1
    g_plain = _graph_of(lambda: per_field(False))
    ms_plain_graph, _ = time_steps(lambda k: g_plain.replay(), pf_steps)
    per_field(True)
    ms_fused_host, _ = time_steps(lambda k: per_field(True), pf_steps)
    g_fused = _graph_of(lambda: per_field(True))
    ms_fused_graph, _ = time_steps(lambda k: g_fused.replay(), pf_steps)
    # batched decompress of the same archives (eager launches), every 16th field's floats
    # compared with its single-field decompress
    dec_out = [torch.empty(max(t.numel(), 4), dtype=torch.float32, device="cuda") for t in d]
    ebx2 = [x["ebx2"] for x in res]
    plan.decompress_async(out.payloads, out.tails, ebx2, dec_out)
    ms_dec, _ = time_steps(lambda k: plan.decompress_async(out.payloads, out.tails, ebx2, dec_out),
                           max(3, steps // 2))
    dec_mismatch = 0
    for i in range(0, nfields, 16):
        r_i = res[i]
        one = plans[i].decompress(out.payloads[i][: r_i["payload_bytes"]], out.tails[i][: r_i["tail_len"]],
                                  r_i["ebx2"])
        dec_mismatch += 0 if torch.equal(one.view(torch.int32), dec_out[i][: one.numel()].view(torch.int32)) else 1
    del dec_out
    gbps = lambda ms: round(in_bytes / (ms * 1e-3) / 1e9, 2)  # noqa: E731
    names = ["range", "range_final", "tile_1d", "tile_2d", "tile_3d", "scan", "gather"]
    r = {
        "fields": nfields, "input_GB": round(in_bytes / 1e9, 3), "CR": round(in_bytes / pay_bytes, 4),
        "batched_graph_ms": round(ms_batch, 4), "batched_GBps": gbps(ms_batch),
        "batched_alg_frac": round((2 * in_bytes + pay_bytes) / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
        "per_field_plain_host_GBps": gbps(ms_plain_host), "per_field_plain_graph_GBps": gbps(ms_plain_graph),
        "per_field_fused_host_GBps": gbps(ms_fused_host), "per_field_fused_graph_GBps": gbps(ms_fused_graph),
        "speedup_vs_plain_host": round(ms_plain_host / ms_batch, 3),
        "speedup_vs_plain_graph": round(ms_plain_graph / ms_batch, 3),
        "speedup_vs_fused_graph": round(ms_fused_graph / ms_batch, 3),
        "kernel_ms": {k: round(v, 5) for k, v in zip(names, kt)},
        "tuning": tune, "groups_sweep": sweep, "bytes_mismatched_fields": mismatch,
        "batched_decompress_ms": round(ms_dec, 4), "batched_decompress_GBps": gbps(ms_dec),
        "decompress_mismatched_fields": dec_mismatch,
    }
    if host_out is not None:  # host copies of the fields for the CPU baseline's batch loop
        host_out.extend((t.cpu().numpy(), s, e) for t, s, e in zip(d, shapes, ebs))
    del g, g_plain, g_fused, plans, outs, out, plan, d
    torch.cuda.empty_cache()
    return r


def bench_serving(rates=(50, 100, 200, 500, 1000, 2000, 20000), duration_s=1.0, pool_fields=48, min_requests=150):
    """Poisson request streams (80% compress / 20% decompress of archives already served)
    over a device-resident pool of fields log-uniform in 256 KB - 1 GB, under a 200 GB HBM
    budget; latency is arrival to completion.  Each rate runs for max(duration_s,
    min_requests / rate) seconds of arrivals, so every p99 rests on >= 100 samples; the last
    rate offers more work than the device serves (~2 TB/s of input at a ~130 MB mean field),
    so requests queue and share batches."""
    import torch

    from fzgpu import serve as S
    t0 = time.perf_counter()
    pool = S.field_pool(pool_fields)
    torch.cuda.synchronize()
    log(f"  pool of {pool_fields} fields ({sum(4 * x.numel() for _, x in pool) / 1e9:.2f} GB) in "
        f"{time.perf_counter() - t0:.1f} s")
    out = {"pool_GB": round(sum(4 * x.numel() for _, x in pool) / 1e9, 3), "duration_s": duration_s}
    # one long-lived device backend (its plan pool holds up to 16 GB of workspaces, as a
    # caching allocator would); warm-up: every pool field compressed and decompressed alone,
    # so that the caching allocator and the plan pool hold every request size before the
    # timed streams
    # workspaces and kept archives come from one 160 GB device arena: no request-stream
    # allocation reaches the caching allocator (decompress outputs, which go to the client, do)
    be = S.DeviceBackend(plan_pool_bytes=16 << 30, arena_bytes=int(160e9))
    warm = S.Server(be, hbm_budget_bytes=200e9, slots=3, max_fields=1, max_archives=len(pool))
    warm.run([S.Request(i, "compress", d, 1e-3, x, 0.0) for i, (d, x) in enumerate(pool)], realtime=False)
    arch = [r.result for r, _ in warm.archives]
    warm2 = S.Server(be, hbm_budget_bytes=200e9, slots=3, max_fields=1)
    warm2.run([S.Request(i, "decompress", a.dims, 1e-3, a, 0.0) for i, a in enumerate(arch)], realtime=False)
    warm2.close()
    warm.close()
    del warm, warm2, arch
    def release(r):
        # the bench reads only the timing of a finished request: a decompress request lets go of
        # its output and of the archive it read, as a client that took its result would, so
        # that device memory held outside the server stays bounded
        if r.kind == "decompress":
            r.result = r.source = None

    for k, rate in enumerate(rates):
        reqs = S.make_requests(rate, max(duration_s, min_requests / rate) if rate < 10000 else 0.25, pool,
                               seed=3000 + k)
        srv = S.Server(be, hbm_budget_bytes=200e9, slots=3, max_fields=256)
        srv.run(reqs, realtime=True, on_arrival=S.pick_archive(4000 + k), on_complete=release)
        out[f"rate_{rate}"] = S.summarize(srv, reqs)
        out[f"rate_{rate}"]["arena_peak_GB"] = round(be.arena.peak / 1e9, 3)
        log(f"  rate {rate}/s: {out[f'rate_{rate}']['latency']} batches {out[f'rate_{rate}']['batch_size_hist']}")
        srv.close()
        del srv, reqs
        torch.cuda.empty_cache()
    out["plan_pool_GB"] = round(be.pool_bytes / 1e9, 3)
    out["arena_GB"] = round(be.arena.size / 1e9, 3)
    out["arena_misses"] = be.arena.misses
    return out


def bench_outlier(x, dims, eb, steps, rotate_bytes=1.2e9):
    """Outlier mode (FIXED payload + side channel) on the same field, inputs rotated as in
    bench_field: input GB/s and the channel's size."""
    import ctypes

    import numpy as np
    import torch

    import fzgpu
    n = x.size
    ncopies = max(2, min(64, math.ceil(rotate_bytes / (4 * n))))
    base = torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
    copies = [base] + [base.clone() for _ in range(ncopies - 1)]
    plan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
    payload, tail, res = plan.alloc_outputs()
    cap = plan.sizes.side_max_bytes
    side = torch.zeros(cap, dtype=torch.uint8, device="cuda")
    L = fzgpu.lib()

    def step(k):
        fzgpu._check(L.fz_compress_outlier_f32(plan.handle, fzgpu._ptr(copies[k % ncopies]), float(eb),
                                               fzgpu._ptr(plan.ws), plan.ws.numel(), fzgpu._ptr(payload),
                                               fzgpu._ptr(tail), fzgpu._ptr(side), cap, fzgpu._ptr(res),
                                               fzgpu._stream()), "fz_compress_outlier")

    for k in range(3):
        step(k)
    ms, _ = time_steps(step, steps)
    r = fzgpu.read_result(res)
    nd, nr = (int(v) for v in side[:16].cpu().view(torch.int64).tolist())
    del copies, base
    torch.cuda.empty_cache()
    return {"GBps": round(4 * n / (ms * 1e-3) / 1e9, 2), "ms": round(ms, 5), "status": r["status"],
            "payload_bytes": r["payload_bytes"], "side_bytes": 16 + 8 * (nd + nr), "n_delta": nd, "n_raw": nr}


def bench_verify(reps=20, sizes=(3600 * 1800, 1 << 24, 1 << 26, 1 << 28)):
    """fz_verify_f32 (verify.hh:25-83 on the device) on reconstruction-like pairs of several
    sizes: microseconds per call and GB/s of the 8n bytes it reads."""
    import torch

    import fzgpu
    g = torch.Generator(device="cuda").manual_seed(7)
    res = {}
    for n in sizes:
        o = torch.randn(n, device="cuda", generator=g).cumsum(0)
        x = o + torch.rand(n, device="cuda", generator=g) * 1e-3
        st = torch.zeros(160, dtype=torch.uint8, device="cuda")
        temp = fzgpu.verify_async(x, o, st)
        for _ in range(3):
            fzgpu.verify_async(x, o, st, temp)
        ms, _ = time_steps(lambda k: fzgpu.verify_async(x, o, st, temp), reps)
        res[str(n)] = {"us": round(ms * 1e3, 2), "GBps": round(8 * n / (ms * 1e-3) / 1e9, 1)}
        log(f"verify n={n}: {ms * 1e3:.1f} us, {8 * n / (ms * 1e-3) / 1e9:.0f} GB/s")
        del o, x
        torch.cuda.empty_cache()
    return res


def bench_staged(x, dims, eb, nreq=24, depth=3, nbuf=4):
    """Host-resident fields (PCIe-inclusive): the reference's pattern (blocking pageable copy in,
    compress, blocking copy of the result out, one field after the other) against the stager
    (pinned inputs, depth requests in flight).  Input GB/s over nreq requests rotating over nbuf
    host copies, every payload checked against the device path."""
    import numpy as np
    import torch

    import fzgpu
    n = x.size
    variants = [(x * np.float32(1 + 0.01 * k)).astype(np.float32) for k in range(nbuf)]
    plan = fzgpu.Plan(*dims)
    ref = []
    for v in variants:
        p, t, r = plan.compress(torch.from_numpy(v).to("cuda"), eb)
        ref.append(p.cpu().numpy())
    # reference pattern: pageable numpy -> device (blocking), compress, record + payload back
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(nreq):
        d = torch.from_numpy(variants[i % nbuf]).to("cuda")
        p, t, r = plan.compress(d, eb)
        out = p.cpu().numpy()
    torch.cuda.synchronize()
    blocking_s = time.perf_counter() - t0
    assert np.array_equal(out, ref[(nreq - 1) % nbuf])
    # stager: pinned inputs, depth in flight
    bufs = []
    for v in variants:
        b = fzgpu.HostBuffer(n, np.float32)
        b.array[:] = v
        bufs.append(b)
    st = fzgpu.Stager(*dims, depth=depth)
    for k in range(depth):  # warm-up
        st.wait(st.compress(bufs[k % nbuf].array, eb))
    t0 = time.perf_counter()
    tickets, mism = [], 0
    for i in range(nreq):
        tickets.append(st.compress(bufs[i % nbuf].array, eb))
        if len(tickets) >= depth:
            k = tickets[len(tickets) - depth]
            res, payload, _ = st.wait(k)
            mism += 0 if np.array_equal(payload, ref[(k - depth) % nbuf]) else 1
    for k in tickets[len(tickets) - depth + 1:]:
        res, payload, _ = st.wait(k)
        mism += 0 if np.array_equal(payload, ref[(k - depth) % nbuf]) else 1
    staged_s = time.perf_counter() - t0
    # the copy-in alone: pinned H2D of the same bytes, back to back
    dev = torch.empty(n, dtype=torch.float32, device="cuda")
    pin = torch.from_numpy(bufs[0].array)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(8):
        dev.copy_(pin, non_blocking=True)
    torch.cuda.synchronize()
    h2d_gbps = 8 * 4 * n / (time.perf_counter() - t0) / 1e9
    st.close()
    res = {"requests": nreq, "depth": depth, "field_MB": round(4 * n / 1e6, 1),
           "blocking_pageable_GBps": round(nreq * 4 * n / blocking_s / 1e9, 2),
           "staged_pinned_GBps": round(nreq * 4 * n / staged_s / 1e9, 2),
           "pinned_h2d_copy_GBps": round(h2d_gbps, 2), "payload_mismatches": mism}
    log(f"staged: {res}")
    return res


def _host_info():
    import platform
    import subprocess
    model = platform.processor() or ""
    try:
        for ln in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if ln.lower().startswith("model name"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": AFFINITY_AT_START, "lscpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def _median_time(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def cpu_baseline(workload, batch_fields=None, reps=5):
    """The CPU side of the comparison: the oracle (a single-threaded restatement of the
    reference kernels) and the same code under OpenMP on the host's cores, compress (range +
    dual-quant + bitshuffle / flags + scan + encode) and decompress, no file I/O; the median of
    `reps` timed runs after one warm-up, on the three single-field workloads (3D at both error
    bounds) and on the uneven batch run as a loop over its fields.  batch_fields: host arrays
    [(x, dims, eb)] of the batch (None skips it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O
    from fzgpu import datagen as D
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    table = {}

    def one(x, dims, eb, nt):
        r = O.compress(x, dims, eb, O.MODE_FIXED, nthreads=nt)
        assert r["status"] == 0
        return r

    for name in D.CONFIGS:
        x, dims, eb0 = D.config_field(name)
        x = np.ascontiguousarray(x.ravel())
        for eb in ([eb0, 1e-4] if dims[2] > 1 else [eb0]):
            row = {}
            ref = one(x, dims, eb, threads)
            rc, xo = O.decompress(ref["payload"], ref["tail"], x.size, dims, ref["info"]["ebx2"], nthreads=threads)
            assert rc == 0
            row["CR"] = round(4 * x.size / ref["info"]["payload_bytes"], 4)
            row["max_err_over_range"] = float(np.abs(xo.astype(np.float64) - x).max() /
                                              (float(x.max()) - float(x.min())))
            for nt, tag in ((1, "1t"), (threads, "omp")):
                tc = _median_time(lambda: one(x, dims, eb, nt), reps)
                td = _median_time(lambda: O.decompress(ref["payload"], ref["tail"], x.size, dims,
                                                       ref["info"]["ebx2"], nthreads=nt), reps)
                row[f"compress_{tag}_GBps"] = round(4 * x.size / tc / 1e9, 4)
                row[f"decompress_{tag}_GBps"] = round(4 * x.size / td / 1e9, 4)
            table[f"{name}@{eb:g}"] = row
            log(f"  cpu {name}@{eb:g}: {row}")
    if batch_fields:
        arch = [one(x, dims, eb, threads) for x, dims, eb in batch_fields]
        nbytes = sum(4 * x.size for x, _, _ in batch_fields)

        def loop_c(nt, fields):
            for x, dims, eb in fields:
                one(x, dims, eb, nt)

        def loop_d(nt, idx):
            for i in idx:
                x, dims, _ = batch_fields[i]
                a = arch[i]
                O.decompress(a["payload"], a["tail"], x.size, dims, a["info"]["ebx2"], nthreads=nt)

        breps = max(1, reps // 2)
        row = {"fields": len(batch_fields), "input_GB": round(nbytes / 1e9, 3),
               "compress_omp_GBps": round(nbytes / _median_time(lambda: loop_c(threads, batch_fields), breps) / 1e9, 4),
               "decompress_omp_GBps": round(nbytes / _median_time(lambda: loop_d(threads, range(len(arch))), breps) /
                                            1e9, 4)}
        sub = list(range(0, len(batch_fields), 16))  # 1-thread: every 16th field, one timed pass
        sb = sum(4 * batch_fields[i][0].size for i in sub)
        row["compress_1t_GBps"] = round(sb / _median_time(lambda: loop_c(1, [batch_fields[i] for i in sub]), 1, 0) /
                                        1e9, 4)
        row["decompress_1t_GBps"] = round(sb / _median_time(lambda: loop_d(1, sub), 1, 0) / 1e9, 4)
        row["sample"] = (f"OpenMP: all fields, median of {breps} after 1 warm-up; 1 thread: every 16th field "
                         f"({len(sub)} fields, {sb / 1e9:.2f} GB), one pass")
        table["uneven-batch-256"] = row
        log(f"  cpu batch: {row}")
    head = table[f"{workload}@{D.config_field(workload)[2]:g}"]
    return dict(value=head["compress_omp_GBps"], unit="GB/s", cores=threads, kind="port",
                sample=(f"oracle compress of the whole {workload} field under OpenMP, {threads} threads, median "
                        f"of {reps} after 1 warm-up; per-workload 1-thread / OpenMP compress and decompress in "
                        f"'table'"),
                host=_host_info(), table=table)


def pmc_target(D):
    """Counter-collection workload: calibration copy (256 MiB read + 256 MiB write through
    hipMemcpyAsync's copy kernel), then 3 fused compress calls on each workload."""
    import torch

    import fzgpu
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    for name in D.CONFIGS:
        x, dims, eb = D.config_field(name)
        d = torch.from_numpy(x.ravel()).cuda()
        plan = fzgpu.Plan(*dims)
        payload, tail, res = plan.alloc_outputs()
        for _ in range(3):
            plan.compress_async(d, eb, payload, tail, res)
        r = fzgpu.read_result(res)
        print(json.dumps({"workload": name, "payload_bytes": r["payload_bytes"], "n": int(x.size)}), flush=True)
        del d


def inverse_target(D, workload):
    """Profiling workload for the inverse and the outlier mode: the calibration copy, then 5
    decompress calls, 5 outlier-mode compress calls and 5 outlier-mode decompress calls on
    one workload (inputs device-resident)."""
    import torch

    import fzgpu
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    x, dims, eb = D.config_field(workload)
    d = torch.from_numpy(x.ravel()).cuda()
    plan = fzgpu.Plan(*dims)
    payload, tail, r = plan.compress(d, eb)
    out = torch.empty_like(d)
    for _ in range(5):
        plan.decompress(payload, tail, r["ebx2"], out=out)
    oplan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
    for _ in range(5):
        op, ot, orr, side = oplan.compress_outlier(d, eb)
    for _ in range(5):
        oplan.decompress_outlier(op, ot, side, orr["ebx2"], orr["vmin"], out=out)
    torch.cuda.synchronize()
    print(json.dumps({"workload": workload, "n": int(x.size), "payload_bytes": r["payload_bytes"],
                      "outlier_payload_bytes": orr["payload_bytes"], "side_bytes": int(side.numel())}), flush=True)


def batch_target(D):
    """Counter / trace target for the batched kernels: 3 compress calls of the 256-field
    uneven batch at the default tuning (one group on one stream), then 3 batched decompress
    calls of its archives."""
    import torch

    import fzgpu
    spec = D.batch_fields()
    d = [D.generate_gpu(f["kind"], f["dims"], f["seed"]) for f in spec]
    plan = fzgpu.BatchPlan([f["dims"] for f in spec])
    out = plan.alloc_outputs()
    for _ in range(3):
        plan.compress_async(d, [f["eb"] for f in spec], out)
    torch.cuda.synchronize()
    res = fzgpu.read_results(out.res, len(spec))
    outs = [torch.empty(max(t.numel(), 4), dtype=torch.float32, device="cuda") for t in d]
    for _ in range(3):
        plan.decompress_async(out.payloads, out.tails, [r["ebx2"] for r in res], outs)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "uneven-batch-256", "input_bytes": sum(4 * t.numel() for t in d),
                      "payload_bytes": sum(r["payload_bytes"] for r in res)}), flush=True)


def load_traffic(kernel, workload):
    p = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("traffic_bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cesm-like-3600x1800")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed runs per CPU-baseline figure (median)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-field uneven batch")
    ap.add_argument("--batch-only", action="store_true", help="only the 256-field uneven batch (prints its JSON)")
    ap.add_argument("--batch-fields", type=int, default=256)
    ap.add_argument("--batch-target", action="store_true",
                    help="profiling target: the 256-field uneven batch, default tuning (one group), 3 compress "
                         "calls, then 3 batched decompress calls")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-serving leg")
    ap.add_argument("--serve-only", action="store_true", help="only the request-serving leg (prints its JSON)")
    ap.add_argument("--no-staged", action="store_true", help="skip the host-resident (PCIe-inclusive) leg")
    ap.add_argument("--inverse-target", action="store_true",
                    help="profiling workload: decompress and the outlier mode on --workload")
    ap.add_argument("--staged-only", action="store_true",
                    help="only the host-resident (PCIe-inclusive) leg: blocking pageable vs stager (prints its JSON)")
    ap.add_argument("--verify-only", action="store_true",
                    help="only the quality-metric reduction at four sizes (prints its JSON)")
    ap.add_argument("--pmc-target", action="store_true",
                    help="profiling target: a 256 MiB device copy (counter calibration), then 3 fused compress "
                         "calls per workload, no timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (fields are independent; "
              f"the multi-GPU layout is a per-field shard with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch

    from fzgpu import datagen as D
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.pmc_target:
        pmc_target(D)
        return
    if args.batch_only:
        print(json.dumps(bench_batch(args.steps, args.warmup, args.batch_fields)), flush=True)
        return
    if args.batch_target:
        batch_target(D)
        return
    if args.verify_only:
        print(json.dumps(bench_verify(max(args.steps, 5))), flush=True)
        return
    if args.inverse_target:
        inverse_target(D, args.workload)
        return
    if args.staged_only:
        x, dims, eb = D.config_field(args.workload)
        print(json.dumps(bench_staged(x.ravel(), dims, eb)), flush=True)
        return
    if args.serve_only:
        print(json.dumps(bench_serving()), flush=True)
        return
    log(f"generating {args.workload}")
    x, dims, eb = D.config_field(args.workload)
    x = x.ravel()
    log("timing fused compress")
    main_r = bench_field(x, dims, eb, args.steps, args.warmup, fused=True, label=args.workload)
    range_ms, tile_ms, gather_ms = (main_r["kernel_ms"] + [0.0, 0.0, 0.0])[:3]
    n = x.size
    # k_tile_compress reads the field once and writes the bitflags and the nonzero blocks
    alg_bytes = 4 * n + main_r["payload_bytes"]
    achieved = alg_bytes / (tile_ms * 1e-3) / 1e9
    extra = {
        "pipeline_alg_GBps": round((8 * n + main_r["payload_bytes"]) / (main_r["ms"] * 1e-3) / 1e9, 2),
        "pipeline_alg_frac": round((8 * n + main_r["payload_bytes"]) / (main_r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
        "warm_mall_GBps": round(main_r["warm_gbps"], 2),
        "kernel_ms": {"range": round(range_ms, 5), "tile_compress": round(tile_ms, 5),
                      "scan_gather": round(gather_ms, 5)},
        "range_kernel_frac": round(4 * n / (range_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
        "decompress_GBps": round(main_r["decompress_gbps"], 2),
        "decompress_kernel_ms": [round(v, 5) for v in main_r["decompress_kernel_ms"]],
        "decompress_verify_ms": round(main_r["decompress_verify_ms"], 5),
        # the last interval is the metrics' two kernels, which read 8n bytes
        "verify_GBps": round(8 * n / (main_r["decompress_verify_kernel_ms"][-1] * 1e-3) / 1e9, 2),
        "verify_PSNR_dB": round(main_r["verify_stats"]["reduced"]["PSNR"], 4),
        "verify_max_rel_err": main_r["verify_stats"]["max_err"]["rel"],
        "input_copies": main_r["copies"],
    }
    log("timing the outlier mode")
    extra["outlier_mode"] = bench_outlier(x, dims, eb, args.steps)
    log("timing plain pipeline")
    plain = bench_field(x, dims, eb, args.steps, args.warmup, fused=False, label="plain")
    extra["plain_pipeline_GBps"] = round(plain["gbps"], 2)
    extra["plain_kernel_ms"] = [round(v, 5) for v in plain["kernel_ms"]]
    if not args.no_extra:
        for name in D.CONFIGS:
            if name == args.workload:
                continue
            log(f"timing {name}")
            xe, de, ebe = D.config_field(name)
            ebs = [ebe, 1e-4] if de[2] > 1 else [ebe]
            for e in ebs:
                r = bench_field(xe.ravel(), de, e, args.steps, args.warmup, fused=True, label=name)
                extra[f"{name}@{e:g}"] = {"GBps": round(r["gbps"], 2), "CR": round(r["cr"], 4),
                                          "fused_kernel_frac": round((4 * xe.size + r["payload_bytes"]) /
                                                                     (r["kernel_ms"][1] * 1e-3) / 1e9 /
                                                                     HBM_PEAK_GBPS, 4),
                                          "decompress_GBps": round(r["decompress_gbps"], 2),
                                          "outlier_mode_GBps": bench_outlier(xe.ravel(), de, e, args.steps)["GBps"]}
    batch_host = [] if not args.no_cpu else None
    if not args.no_batch:
        log(f"timing the {args.batch_fields}-field uneven batch")
        extra["uneven_batch"] = bench_batch(min(args.steps, 20), min(args.warmup, 3), args.batch_fields,
                                            host_out=batch_host)
    if not args.no_serve:
        log("serving loop")
        extra["serving"] = bench_serving()
    if not args.no_staged:
        log("host-resident fields (PCIe-inclusive)")
        extra["host_staged"] = bench_staged(x, dims, eb)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(args.workload, batch_host or None, args.cpu_reps)
        batch_host = None
    line = {
        "metric": METRIC,
        "value": round(main_r["gbps"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded; " + WORKLOAD_TEXT.get(args.workload, args.workload) + ")",
        "config": {"workload": args.workload, "dims_xyz": list(dims), "eb_rel": eb, "mode": "fixed",
                   "pipeline": "fused (range + tile_compress + scan_gather kernels)",
                   "inputs": "HBM-resident, rotated over copies (MALL-cold)"},
        "compression_ratio": round(main_r["cr"], 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": load_traffic("tile_compress", args.workload),
                     "kernel": "k_tile_compress", "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
