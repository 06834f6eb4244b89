"""Request serving on top of the batched C ABI: continuous batching under an HBM budget.

Compress and decompress requests arrive over time.  A request is admitted when its device
memory (outputs and its share of the batch workspace) fits in what the budget has left;
admitted requests of one kind that are waiting together form the next batch, which runs on
its own stream through fz_batch_compress_f32 / fz_batch_decompress_f32 while earlier batches
are still on the device.  When a batch's event completes, its requests finish, its memory
returns to the budget and freed slots admit the next batch.  Inputs and archives stay
device-resident: a compress request reads a field that is already in HBM and leaves its
archive (payload + tail + result record) in HBM for later decompress requests.

The scheduler only talks to a backend (launch a batch, poll it, account memory), so its
admission and ordering rules run on the CPU with a simulated backend in the tests.
"""
import bisect
import collections
import dataclasses
import math
import time

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                 # "compress" | "decompress"
    dims: tuple               # (x, y, z)
    eb: float = 1e-3          # compress: value-range-relative error bound
    source: object = None     # compress: input tensor (device); decompress: the Archive
    arrival: float = 0.0      # seconds, on the server's clock
    admitted: float = math.nan
    done: float = math.nan
    batch: int = -1
    result: object = None     # compress: Archive; decompress: output tensor

    @property
    def n(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    @property
    def input_bytes(self):
        return 4 * self.n

    @property
    def latency(self):
        return self.done - self.arrival


@dataclasses.dataclass
class Archive:
    """A compressed field kept on the device: payload[:payload_bytes], tail[:tail_len]."""
    dims: tuple
    payload: object
    tail: object
    record: dict
    keep: object = None       # the batch buffers the views above live in
    buffer: object = None     # the server's OutputBuffer accounting for `keep`
    side: object = None       # outlier mode: the side channel (its capacity slice)
    pins: int = 0             # waiting or running decompress requests that read this archive

    @property
    def nbytes(self):
        return int(self.record["payload_bytes"]) + 2 * int(self.record["tail_len"])


class OutputBuffer:
    """Device outputs of one finished compress batch.  Its archives are views into one
    allocation, so its bytes stay in the budget until the last of them is dropped; then
    release() (if set) gives the memory back to the backend that allocated it."""

    def __init__(self, nbytes, release=None):
        self.nbytes = int(nbytes)
        self.refs = 0
        self.release = release


class HbmBudget:
    """Byte accounting of the device memory the server may hold."""

    def __init__(self, total_bytes):
        self.total = int(total_bytes)
        self.used = 0
        self.peak = 0

    def fits(self, nbytes):
        return self.used + nbytes <= self.total

    def take(self, nbytes):
        assert self.fits(nbytes)
        self.used += int(nbytes)
        self.peak = max(self.peak, self.used)

    def give(self, nbytes):
        self.used -= int(nbytes)
        assert self.used >= 0


def request_bytes(req, side_fraction=0.0):
    """(workspace, output) device bytes of a request, upper bounds: its share of the batch
    workspace (staging / codes, 4096 B per tile, plus counters) and its output (the archive
    slot of a compress request, the floats of a decompress request).  In the outlier mode
    (side_fraction > 0) an archive also reserves that fraction of its largest side channel
    (16 + 8 n bytes), and the workspace is the single-field outlier one (about 18 n bytes)."""
    n = req.n
    tiles = n // 2048 + 1
    w = (n + 255) // 256
    ws = 4096 * tiles + 32 * tiles + 1024
    out = (4 * w + 4096 * tiles + 4096 + 64) if req.kind == "compress" else 4 * n + 64
    if side_fraction > 0:
        ws = 18 * n + 16 * tiles + 65536
        if req.kind == "compress":
            out += 16 + 8 * math.ceil(n * side_fraction) + 256
    return ws, out


class Server:
    """Continuous-batching scheduler.

    backend: object with
        launch(kind, requests) -> handle     (allocates outputs, enqueues the batch)
        poll(handle) -> bool                 (True when the batch finished on the device)
        finish(handle, requests) -> None     (fills request.result)
    slots: batches allowed on the device at once; max_fields: fields per batch.
    """

    def __init__(self, backend, hbm_budget_bytes=200e9, slots=3, max_fields=256, max_archives=64,
                 clock=time.perf_counter, sleep=time.sleep):
        self.backend = backend
        # the outlier mode reserves side-channel room per archive (DeviceBackend(mode=OUTLIER))
        self.side_fraction = float(getattr(backend, "side_fraction", 0.0))
        self.budget = HbmBudget(hbm_budget_bytes)
        self.max_archives = int(max_archives)
        self.archives = collections.deque()  # (request, OutputBuffer) of compress results kept on the device
        self.slots = int(slots)
        self.max_fields = int(max_fields)
        self.clock = clock
        self.sleep = sleep
        self.t0 = None
        self.waiting = {"compress": collections.deque(), "decompress": collections.deque()}
        self.inflight = []        # (handle, kind, requests, bytes)
        self.completed = []
        self.batches = []         # (batch id, kind, n requests, input bytes, launch time, finish time)
        self.max_inflight = 0
        self.budget_limited = 0   # admission attempts stopped because the next request did not fit the budget
        self.max_need = 0         # the largest request submitted so far (device bytes)
        self.on_complete = None

    def now(self):
        return self.clock() - self.t0

    def submit(self, req):
        need = sum(self._bytes(req))
        if need > self.budget.total:
            raise ValueError(f"request {req.rid} needs more device memory than the whole budget")
        self.max_need = max(self.max_need, need)
        if req.kind == "decompress" and isinstance(req.source, Archive):
            req.source.pins += 1  # not evicted while this request waits or runs
        self.waiting[req.kind].append(req)

    def _bytes(self, req):
        return request_bytes(req, self.side_fraction)

    def _pinned_bytes(self):
        return sum({id(buf): buf.nbytes for req, buf in self.archives
                    if req.result is not None and req.result.pins > 0}.values())

    def _form(self, kind):
        """FIFO batch of waiting requests of `kind` that fits the budget (a request that does
        not fit blocks the ones behind it, so admission order is arrival order)."""
        q = self.waiting[kind]
        take, nbytes = [], 0
        while q and len(take) < self.max_fields:
            need = sum(self._bytes(q[0]))
            # archives are dropped only when dropping them makes the request fit; memory held
            # by running batches is not theirs to free
            if not self.budget.fits(nbytes + need) and \
                    self.budget.used - self._freeable() + nbytes + need <= self.budget.total:
                while not self.budget.fits(nbytes + need) and self._evict():
                    pass
            if not self.budget.fits(nbytes + need):
                self.budget_limited += 1
                break
            take.append(q.popleft())
            nbytes += need
        return take, nbytes

    def _freeable(self):
        """Bytes that evicting every archive no decompress request is reading would return."""
        blocked = {id(buf) for req, buf in self.archives if req.result is not None and req.result.pins > 0}
        return sum({id(buf): buf.nbytes for _, buf in self.archives if id(buf) not in blocked}.values())

    def _drop_ref(self, buf):
        buf.refs -= 1
        if buf.refs == 0:
            self.budget.give(buf.nbytes)
            if buf.release is not None:
                buf.release()
                buf.release = None

    def close(self):
        """Drops every archive the server holds (their buffers go back to the backend); a
        server that shares its backend with later servers is closed when it is done."""
        for req, buf in list(self.archives):
            req.result = None
            self._drop_ref(buf)
        self.archives.clear()

    def _evict(self):
        """Drops the oldest archive no decompress request is reading; its batch's output
        buffer returns to the budget with its last archive.  False when nothing can go."""
        for i, (req, buf) in enumerate(self.archives):
            if req.result is not None and req.result.pins > 0:
                continue
            del self.archives[i]
            req.result = None
            self._drop_ref(buf)
            return True
        return False

    def archive_pool(self):
        """Compress results currently kept on the device that a new decompress request may read.
        Pinned archives (read by waiting decompress requests) cannot be evicted, so an archive
        whose buffer is not pinned yet is offered only while the pinned bytes leave room for the
        largest request: the head of every queue then fits once the running batches finish, and
        admission cannot deadlock on memory held for requests that are still waiting."""
        pinned = {id(buf) for req, buf in self.archives if req.result is not None and req.result.pins > 0}
        held = self._pinned_bytes()
        out = []
        for r, buf in self.archives:
            if r.result is None:
                continue
            need = max(self.max_need, sum(self._bytes(Request(-1, "decompress", r.dims))))
            if id(buf) in pinned or held + buf.nbytes + need <= self.budget.total:
                out.append(r)
        return out

    def _launch_ready(self):
        launched = False
        # the kind whose oldest request waited longest goes first
        kinds = sorted((k for k in self.waiting if self.waiting[k]), key=lambda k: self.waiting[k][0].arrival)
        for kind in kinds:
            if len(self.inflight) >= self.slots:
                break
            reqs, nbytes = self._form(kind)
            if not reqs:
                continue
            self.budget.take(nbytes)
            t = self.now()
            bid = len(self.batches)
            for r in reqs:
                r.admitted = t
                r.batch = bid
            handle = self.backend.launch(kind, reqs)
            self.inflight.append((handle, kind, reqs, nbytes, bid))
            self.batches.append([bid, kind, len(reqs), sum(r.input_bytes for r in reqs), t, math.nan])
            self.max_inflight = max(self.max_inflight, len(self.inflight))
            launched = True
        return launched

    def _poll(self):
        done_any = False
        still = []
        for item in self.inflight:
            handle, kind, reqs, nbytes, bid = item
            if self.backend.poll(handle):
                t = self.now()
                self.backend.finish(handle, reqs)
                for r in reqs:
                    r.done = t
                    if self.on_complete:
                        self.on_complete(r)
                self.completed.extend(reqs)
                out_bytes = sum(self._bytes(r)[1] for r in reqs) if kind == "compress" else 0
                self.budget.give(nbytes - out_bytes)  # the workspace share; decompress outputs go to the caller
                if kind == "decompress":
                    for r in reqs:
                        if isinstance(r.source, Archive):
                            r.source.pins -= 1
                elif self.max_archives > 0:  # archives stay on the device, their buffer in the budget
                    keep = reqs[0].result.keep if isinstance(reqs[0].result, Archive) else None
                    free = getattr(self.backend, "free_outputs", None)
                    buf = OutputBuffer(out_bytes, (lambda k=keep: free(k)) if free and keep is not None else None)
                    buf.refs = 1  # held until every archive of the batch is registered
                    for r in reqs:
                        if isinstance(r.result, Archive):
                            r.result.buffer = buf
                            buf.refs += 1
                            self.archives.append((r, buf))
                    self._drop_ref(buf)
                    while len(self.archives) > self.max_archives and self._evict():
                        pass
                else:
                    self.budget.give(out_bytes)
                self.batches[bid][5] = t
                done_any = True
            else:
                still.append(item)
        self.inflight = still
        return done_any

    def run(self, requests, realtime=True, on_arrival=None, on_complete=None):
        """Serve `requests` (arrival times in seconds from the start).  With realtime, a
        request becomes visible when the clock passes its arrival; otherwise all are
        visible at once.  on_arrival(server, req) may complete the request before it queues
        (a decompress request picks its archive there); on_complete(req) sees each finished
        request before its archive can be evicted."""
        self.on_complete = on_complete
        # archives the server keeps are its to free when it drops them; with max_archives = 0
        # they go to the caller, so the backend must not take them from memory it reuses
        self.backend.server_keeps_outputs = self.max_archives > 0
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = self.clock()
        i = 0
        while i < len(pending) or any(self.waiting.values()) or self.inflight:
            t = self.now()
            while i < len(pending) and (not realtime or pending[i].arrival <= t):
                if on_arrival:
                    on_arrival(self, pending[i])
                self.submit(pending[i])
                i += 1
            self._poll()
            launched = self._launch_ready()
            if not launched and not self.inflight and any(self.waiting.values()) and i >= len(pending):
                # nothing runs, nothing arrives and nothing can be admitted: no later step changes that
                raise RuntimeError("admission deadlock: waiting requests can never fit the budget")
            if not self.inflight and not any(self.waiting.values()) and i < len(pending) and realtime:
                # idle until shortly before the next arrival, then poll: a sleeping thread
                # wakes late by tens to hundreds of microseconds, which would land on the
                # request's latency
                dt = pending[i].arrival - self.now()
                if dt > 1.5e-3:
                    self.sleep(dt - 1e-3)
        return self.completed


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs], dtype=np.float64) * 1e3
    if lat.size == 0:
        return {"n": 0}
    # a 99th percentile of fewer than 100 samples is only their maximum: not reported
    p99 = round(float(np.percentile(lat, 99)), 4) if lat.size >= 100 else None
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)), 4),
            "p99_ms": p99, "mean_ms": round(float(lat.mean()), 4),
            "max_ms": round(float(lat.max()), 4)}


def poisson_arrivals(rate, duration_s, rng):
    """Arrival times of a Poisson process of `rate` per second over [0, duration_s)."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            return out
        out.append(t)


# ---------------------------------------------------------------- the device backend

class DeviceArena:
    """One device allocation carved into 256-byte-aligned blocks: best fit from a free list kept
    in address order, a freed block merged with free neighbours.  The serving backend takes its
    batch workspaces and archive buffers from it, so a request stream of ever-new sizes never
    reaches the caching allocator (whose splitting of cached segments fragments them until it
    must free them all, a device-wide wait).  With buf=None only the offsets are managed (the
    CPU tests)."""

    ALIGN = 256

    def __init__(self, nbytes, device="cuda", buf=None):
        self.size = int(nbytes) // self.ALIGN * self.ALIGN
        if buf is None and device is not None:
            import torch
            buf = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.buf = buf
        self.free = [(0, self.size)]  # (offset, bytes), by offset, never two adjacent
        self.used = 0
        self.peak = 0
        self.misses = 0               # requests no free block could hold (counted by the caller)

    def alloc(self, nbytes):
        """(offset, bytes) of a block of at least nbytes, or None when no free block holds it."""
        n = max(self.ALIGN, (int(nbytes) + self.ALIGN - 1) // self.ALIGN * self.ALIGN)
        best = -1
        for i, (o, b) in enumerate(self.free):
            if b >= n and (best < 0 or b < self.free[best][1]):
                best = i
                if b == n:
                    break
        if best < 0:
            return None
        o, b = self.free[best]
        if b == n:
            del self.free[best]
        else:
            self.free[best] = (o + n, b - n)
        self.used += n
        self.peak = max(self.peak, self.used)
        return (o, n)

    def release(self, block):
        o, n = block
        i = bisect.bisect_left(self.free, (o, 0))
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] > o or i < len(self.free) and o + n > self.free[i][0]:
            raise ValueError("block released twice or not from this arena")
        self.used -= n
        if i < len(self.free) and o + n == self.free[i][0]:  # merge with the next free block
            n += self.free[i][1]
            del self.free[i]
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == o:  # and with the previous one
            po, pn = self.free[i - 1]
            self.free[i - 1] = (po, pn + n)
        else:
            self.free.insert(i, (o, n))

    def tensor(self, block, nbytes):
        o, _ = block
        return self.buf[o: o + int(nbytes)]


class PlanPool:
    """Free plans of finished batches keyed by their field shapes: take() hands out a free plan
    of those shapes (or None), give() returns one, dropping the least recently used keys' plans
    while more than max_bytes of workspace would be held (a plan larger than that is dropped
    at once).  nbytes(plan) gives a plan's workspace size and close(plan) releases it."""

    def __init__(self, max_bytes, nbytes=lambda p: p.ws.numel(), close=lambda p: p.close()):
        self.max_bytes = int(max_bytes)
        self.nbytes = nbytes
        self.close = close
        self.free = collections.OrderedDict()  # shapes -> [plans], least recently used first
        self.bytes = 0

    def take(self, shapes):
        lst = self.free.get(shapes)
        if not lst:
            return None
        plan = lst.pop()
        self.bytes -= self.nbytes(plan)
        if not lst:
            del self.free[shapes]
        return plan

    def drop_lru(self):
        """Closes the least recently used free plan; False when none is left."""
        if not self.free:
            return False
        old_shapes, lst = next(iter(self.free.items()))
        old = lst.pop(0)
        self.bytes -= self.nbytes(old)
        self.close(old)
        if not lst:
            del self.free[old_shapes]
        return True

    def give(self, shapes, plan):
        nb = self.nbytes(plan)
        if nb > self.max_bytes:
            self.close(plan)
            return
        while self.free and self.bytes + nb > self.max_bytes:
            old_shapes, lst = next(iter(self.free.items()))
            old = lst.pop(0)
            self.bytes -= self.nbytes(old)
            self.close(old)
            if not lst:
                del self.free[old_shapes]
        self.free.setdefault(shapes, []).append(plan)
        self.free.move_to_end(shapes)
        self.bytes += nb


class DeviceBackend:
    """Batches through the batched C ABI, one stream per batch from a small stream pool.

    The plans of finished batches are kept for reuse in a PlanPool (at most `plan_pool_bytes`
    of workspace held, like blocks kept by a caching allocator), so a batch of a shape seen
    before skips plan creation and workspace set-up.  A compress batch's result records are
    copied into pinned host memory behind its kernels, so finishing the batch reads host
    memory and waits for nothing."""

    def __init__(self, mode=0, nstreams=3, device="cuda", plan_pool_bytes=4 << 30, side_fraction=1.0,
                 arena_bytes=0):
        import torch
        import fzgpu
        self.torch = torch
        self.mode = mode
        # outlier mode: archives carry a side channel with room for this fraction of values
        self.side_fraction = float(side_fraction) if mode == fzgpu.MODE_OUTLIER else 0.0
        self.device = device
        self.streams = [torch.cuda.Stream(device=device) for _ in range(nstreams)]
        self.busy = [False] * nstreams
        self.pool = PlanPool(plan_pool_bytes, close=self._close_plan)
        # workspaces and kept archives from one arena (arena_bytes > 0), else from torch
        self.arena = DeviceArena(arena_bytes, device) if arena_bytes else None
        self.server_keeps_outputs = False  # set by the server: it frees the archives it keeps

    @property
    def pool_bytes(self):
        return self.pool.bytes

    def _block(self, nbytes, blocks):
        """A uint8 tensor of nbytes from the arena (its block appended to blocks), or None.  The
        free plans' workspaces give way first, least recently used first."""
        if self.arena is None:
            return None
        b = self.arena.alloc(nbytes)
        while b is None and self.pool.drop_lru():
            b = self.arena.alloc(nbytes)
        if b is None:
            self.arena.misses += 1
            return None
        blocks.append(b)
        return self.arena.tensor(b, nbytes)

    def _close_plan(self, plan):
        for b in getattr(plan, "arena_blocks", ()):
            self.arena.release(b)
        plan.arena_blocks = []
        plan.close()

    def free_outputs(self, out):
        """Returns a kept batch's archive buffers to the arena (called by the server when it
        drops the batch's last archive)."""
        for b in getattr(out, "arena_blocks", ()):
            self.arena.release(b)
        out.arena_blocks = []

    def _take_plan(self, shapes):
        import fzgpu
        plan = self.pool.take(shapes)
        # a new plan sets up its workspace on the current stream
        if plan is None:
            plan = fzgpu.BatchPlan(list(shapes), mode=self.mode, side_fraction=self.side_fraction or 1.0,
                                   workspace=False)
            plan.arena_blocks = []
            ws = self._block(plan.sizes.workspace_bytes, plan.arena_blocks)
            plan.attach_workspace(ws if ws is not None else fzgpu._empty(fzgpu._bucket(plan.sizes.workspace_bytes)))
        return plan

    def launch(self, kind, reqs):
        torch = self.torch
        si = self.busy.index(False)
        self.busy[si] = True
        s = self.streams[si]
        shapes = tuple(tuple(r.dims) for r in reqs)
        res_host = None
        with torch.cuda.stream(s):
            plan = self._take_plan(shapes)
            if kind == "compress":
                blocks = []
                keep_in_arena = self.arena is not None and self.server_keeps_outputs
                out = plan.alloc_outputs((lambda n: self._block(n, blocks)) if keep_in_arena else None)
                out.arena_blocks = blocks
                if self.side_fraction:
                    plan.compress_outlier_async([r.source for r in reqs], [r.eb for r in reqs], out, stream=s)
                else:
                    plan.compress_async([r.source for r in reqs], [r.eb for r in reqs], out, stream=s)
                res_host = torch.empty(out.res.numel(), dtype=torch.uint8, pin_memory=True)
                res_host.copy_(out.res, non_blocking=True)
                keep = out
            else:
                pays = [r.source.payload for r in reqs]
                tails = [r.source.tail for r in reqs]
                outs = [torch.empty(max(r.n, 4), dtype=torch.float32, device=self.device) for r in reqs]
                ebx2 = [r.source.record["ebx2"] for r in reqs]
                if self.side_fraction:  # side channels by capacity: the kernels read their counts
                    plan.decompress_outlier_async(pays, tails, [r.source.side for r in reqs], ebx2,
                                                  [r.source.record["vmin"] for r in reqs], outs, stream=s)
                else:
                    plan.decompress_async(pays, tails, ebx2, outs, stream=s)
                keep = outs
            ev = torch.cuda.Event()
            ev.record(s)
        return {"si": si, "ev": ev, "keep": keep, "kind": kind, "plan": plan, "shapes": shapes, "res": res_host}

    def poll(self, h):
        return h["ev"].query()

    def finish(self, h, reqs):
        import fzgpu
        self.busy[h["si"]] = False
        self.pool.give(h["shapes"], h["plan"])
        out = h["keep"]
        if h["kind"] == "compress":
            recs = fzgpu.read_results(h["res"], len(reqs))
            for i, r in enumerate(reqs):
                rec = recs[i]
                r.result = Archive(r.dims, out.payloads[i][: rec["payload_bytes"]], out.tails[i][: rec["tail_len"]],
                                   rec, keep=out, side=out.sides[i] if out.sides else None)
        else:
            for i, r in enumerate(reqs):
                r.result = out[i][: r.n]


# ---------------------------------------------------------------- the serving workload

SERVE_EBS = (1e-2, 1e-3, 1e-4)


def field_pool(nfields=48, lo=2 ** 16, hi=2 ** 28, seed=3000, device="cuda"):
    """Device-resident input fields: sizes log-uniform in [lo, hi] floats (256 KB - 1 GB),
    1D / 2D / 3D shapes in turn, content from the device generators."""
    from . import datagen as D
    rng = np.random.default_rng(seed)
    lr = (math.log(1 / 8), math.log(8))
    pool = []
    for i in range(nfields):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        nd = 1 + i % 3
        if nd == 1:
            dims = (n, 1, 1)
        elif nd == 2:
            y = max(2, int(round(math.sqrt(n / math.exp(rng.uniform(*lr))))))
            dims = (max(2, n // y), y, 1)
        else:
            a1, a2 = math.exp(rng.uniform(*lr)), math.exp(rng.uniform(*lr))
            z = max(2, int(round((n / (a1 * a2 * a2)) ** (1 / 3))))
            y = max(2, int(round(a2 * z)))
            dims = (max(2, n // (y * z)), y, z)
        pool.append((dims, D.generate_gpu(D.BATCH_KINDS[i % 3], dims, seed + i, device)))
    return pool


def make_requests(rate, duration_s, pool, seed, decompress_share=0.2):
    """Poisson arrivals; each is a decompress request with probability decompress_share
    (its archive is picked when it arrives), else a compress request of a pool field."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i, t in enumerate(poisson_arrivals(rate, duration_s, rng)):
        dims, x = pool[int(rng.integers(0, len(pool)))]
        eb = SERVE_EBS[int(rng.integers(0, len(SERVE_EBS)))]
        kind = "decompress" if rng.random() < decompress_share else "compress"
        reqs.append(Request(i, kind, dims, eb, x, t))
    return reqs


def pick_archive(seed):
    """on_arrival hook: a decompress request reads a random archive the server holds (a
    compress request of its field when there is none yet)."""
    rng = np.random.default_rng(seed)

    def hook(server, req):
        if req.kind != "decompress":
            return
        pool = [r for r in server.archive_pool() if r.result is not None]
        if not pool:
            req.kind = "compress"
            return
        src = pool[int(rng.integers(0, len(pool)))]
        req.dims = src.dims
        req.source = src.result
    return hook


def summarize(server, reqs):
    done = [r for r in reqs if not math.isnan(r.done)]
    t_end = max(r.done for r in done) if done else 0.0
    t_start = min(r.arrival for r in done) if done else 0.0
    nbytes = sum(r.input_bytes for r in done)
    sizes = [b[2] for b in server.batches]
    hist = collections.Counter(sizes)
    return {
        "requests": len(done), "compress": sum(r.kind == "compress" for r in done),
        "decompress": sum(r.kind == "decompress" for r in done),
        "latency": latency_stats(done),
        "latency_compress": latency_stats([r for r in done if r.kind == "compress"]),
        "latency_decompress": latency_stats([r for r in done if r.kind == "decompress"]),
        "GBps": round(nbytes / max(t_end - t_start, 1e-9) / 1e9, 2),
        "batches": len(server.batches), "mean_batch": round(float(np.mean(sizes)), 2) if sizes else 0,
        "max_batch": max(sizes) if sizes else 0,
        "batch_size_hist": {str(k): hist[k] for k in sorted(hist)},
        "budget_blocked_admissions": server.budget_limited,
        "max_inflight_batches": server.max_inflight, "hbm_peak_GB": round(server.budget.peak / 1e9, 3),
    }
