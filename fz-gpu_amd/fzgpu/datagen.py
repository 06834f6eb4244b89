"""Seeded synthetic float32 fields for the FZ workloads.

Dimensions follow the reference CLI (fz.cu:231-233): ``dims = (x, y, z)`` with x the
fastest-varying axis; numpy arrays are shaped ``(z, y, x)``.  Every generator is
deterministic in its seed (numpy PCG64).

* ``waves``  -- sum of plane waves (wavelengths log-uniform, amplitude ~ wavelength),
  min-max normalised, N(0, 1e-3) noise, clipped to [0, 1], smooth blobs of exact 0.
* ``peaks``  -- detector-like: exact zeros, Gaussian peaks (sigma 1-3 px, log-uniform
  heights 1e1-1e4, truncated at 3 sigma) and N(0, 0.5) background on a contiguous 40%
  segment of 5% of the rows; peaks are added until the exact-zero fraction is 97%.
* ``blobs``  -- sum of isotropic Gaussians (radii 5-40 cells, sigma = r/2), scaled to a
  maximum of 1e-3, everything below the 70th percentile set to exact 0.
"""
import math

import numpy as np

# workload presets: name -> (dims (x, y, z), generator, seed, eb_rel)
CONFIGS = {
    "cesm-like-3600x1800": ((3600, 1800, 1), "waves", 1001, 1e-3),
    "detector-39200x388": ((39200, 388, 1), "peaks", 1002, 1e-3),
    "blobs3d-500x500x100": ((500, 500, 100), "blobs", 1003, 1e-3),
}


def shape_of(dims):
    x, y, z = dims
    return (z, y, x)


def _squeeze_shape(dims):
    """numpy shape with the singleton slow axes dropped (1D / 2D / 3D by the reference rule)."""
    x, y, z = dims
    if z == 1 and y == 1:
        return (x,)
    if z == 1:
        return (y, x)
    return (z, y, x)


def _plane_waves(shp, rng, nmodes, lam_lo, lam_hi):
    """sum_k lam_k * cos(k . r + phi_k) over a grid, built separably from per-axis phasors."""
    acc = np.zeros(shp, dtype=np.float64)
    nd = len(shp)
    for _ in range(nmodes):
        lam = math.exp(rng.uniform(math.log(lam_lo), math.log(lam_hi)))
        v = rng.normal(size=nd)
        v /= np.linalg.norm(v) + 1e-300
        phi = rng.uniform(0.0, 2.0 * math.pi)
        k = 2.0 * math.pi / lam * v
        term = np.exp(1j * phi) * np.ones((1,) * nd, dtype=np.complex128)
        for ax in range(nd):
            e = np.exp(1j * k[ax] * np.arange(shp[ax], dtype=np.float64))
            sh = [1] * nd
            sh[ax] = shp[ax]
            term = term * e.reshape(sh)
        acc += lam * term.real
    return acc


def _quantile_threshold(a, q):
    flat = a.ravel()
    kth = min(flat.size - 1, int(q * flat.size))
    return np.partition(flat, kth)[kth]


def waves(dims, seed, nmodes=32, lam=(20.0, 900.0), zero_frac=0.15, noise=1e-3):
    rng = np.random.default_rng(seed)
    shp = _squeeze_shape(dims)
    f = _plane_waves(shp, rng, nmodes, lam[0], lam[1])
    f -= f.min()
    f /= max(f.max(), 1e-300)
    f += rng.normal(0.0, noise, size=shp)
    np.clip(f, 0.0, 1.0, out=f)
    if zero_frac > 0:
        m = _plane_waves(shp, rng, 8, 300.0, max(300.0, 2.0 * max(shp)))
        thr = _quantile_threshold(m, zero_frac)
        f[m < thr] = 0.0
    return f.astype(np.float32).reshape(shape_of(dims))


def _add_gaussian(f, centre, sigma, height, radius):
    """add height * exp(-|r - c|^2 / (2 sigma^2)) inside the sphere |r - c| <= radius."""
    nd = f.ndim
    sl, axes = [], []
    for ax in range(nd):
        lo = max(0, int(math.floor(centre[ax] - radius)))
        hi = min(f.shape[ax], int(math.ceil(centre[ax] + radius)) + 1)
        if lo >= hi:
            return
        sl.append(slice(lo, hi))
        axes.append(np.arange(lo, hi, dtype=np.float64) - centre[ax])
    r2 = np.zeros([len(a) for a in axes], dtype=np.float64)
    for ax, a in enumerate(axes):
        sh = [1] * nd
        sh[ax] = len(a)
        r2 = r2 + (a * a).reshape(sh)
    g = height * np.exp(-r2 / (2.0 * sigma * sigma))
    g[r2 > radius * radius] = 0.0
    f[tuple(sl)] += g


def peaks(dims, seed, zero_frac=0.97, bg_rows=0.05, bg_seg=0.40, bg_sigma=0.5):
    rng = np.random.default_rng(seed)
    shp = _squeeze_shape(dims)
    f = np.zeros(shp, dtype=np.float64)
    lines = f.reshape(-1, shp[-1])
    nrows, rowlen = lines.shape
    nbg = max(1, int(round(bg_rows * nrows))) if nrows > 1 else 0
    seg = max(1, int(bg_seg * rowlen))
    for r in (rng.choice(nrows, nbg, replace=False) if nbg else []):
        s = int(rng.integers(0, rowlen - seg + 1))
        lines[r, s:s + seg] += rng.normal(0.0, bg_sigma, seg)
    target_nz = (1.0 - zero_frac) * f.size
    for _ in range(1_000_000):
        deficit = target_nz - np.count_nonzero(f)
        if deficit <= 0:
            break
        for _ in range(max(1, min(16, int(deficit // 400)))):
            c = [rng.uniform(0, n) for n in shp]
            sig = rng.uniform(1.0, 3.0)
            h = math.exp(rng.uniform(math.log(1e1), math.log(1e4)))
            _add_gaussian(f, c, sig, h, 3.0 * sig)
    return f.astype(np.float32).reshape(shape_of(dims))


def blobs(dims, seed, ncentres=200, rad=(5.0, 40.0), vmax=1e-3, zero_frac=0.70):
    rng = np.random.default_rng(seed)
    shp = _squeeze_shape(dims)
    f = np.zeros(shp, dtype=np.float64)
    for _ in range(ncentres):
        c = [rng.uniform(0, n) for n in shp]
        r = rng.uniform(rad[0], rad[1])
        sig = r / 2.0
        _add_gaussian(f, c, sig, 1.0, 3.0 * sig)
    f *= vmax / max(f.max(), 1e-300)
    thr = _quantile_threshold(f, zero_frac)
    f[f < thr] = 0.0
    return f.astype(np.float32).reshape(shape_of(dims))


GENERATORS = {"waves": waves, "peaks": peaks, "blobs": blobs}


def generate(kind, dims, seed, **kw):
    return GENERATORS[kind](dims, seed, **kw)


def config_field(name):
    dims, kind, seed, eb = CONFIGS[name]
    return generate(kind, dims, seed), dims, eb


# ---------------------------------------------------------------- the uneven batch

BATCH_EBS = (1e-2, 1e-3, 1e-4)
BATCH_KINDS = ("waves", "peaks", "blobs")


def batch_fields(nfields=256, lo=2 ** 16, hi=2 ** 26, split=(77, 102, 77), seed=2000):
    """Shapes, content generator, seed and eb of every field of the uneven batch.

    Sizes are log-uniform in [lo, hi] elements; the 1D / 2D / 3D split (scaled to nfields)
    is a seeded permutation; aspect ratios are log-uniform in [1/8, 8] per axis pair; eb is
    uniform over BATCH_EBS; the generator cycles waves / peaks / blobs; field i uses seed
    seed + i.  Returns a list of dicts (dims, kind, seed, eb)."""
    tot = sum(split)
    counts = [int(round(nfields * s / tot)) for s in split]
    counts[1] = nfields - counts[0] - counts[2]
    nds = np.array([1] * counts[0] + [2] * counts[1] + [3] * counts[2])
    np.random.default_rng(seed).shuffle(nds)
    lr = (math.log(1 / 8), math.log(8))
    out = []
    for i in range(nfields):
        r = np.random.default_rng(seed + i)
        n = int(round(math.exp(r.uniform(math.log(lo), math.log(hi)))))
        nd = int(nds[i])
        if nd == 1:
            dims = (n, 1, 1)
        elif nd == 2:
            a = math.exp(r.uniform(*lr))
            y = max(2, int(round(math.sqrt(n / a))))
            dims = (max(2, int(round(n / y))), y, 1)
        else:
            a1, a2 = math.exp(r.uniform(*lr)), math.exp(r.uniform(*lr))
            z = max(2, int(round((n / (a1 * a2 * a2)) ** (1 / 3))))
            y = max(2, int(round(a2 * z)))
            dims = (max(2, int(round(n / (y * z)))), y, z)
        out.append(dict(dims=dims, kind=BATCH_KINDS[i % 3], seed=seed + i, eb=BATCH_EBS[int(r.integers(0, 3))]))
    return out


# ---------------------------------------------------------------- device-side generators
#
# Same families as above, built on the GPU with torch so that multi-GB batches are made in
# seconds: a smooth field is a sum of plane waves evaluated from the flat index in chunks.

def _smooth_gpu(shp, rng, nmodes, lam_lo, lam_hi, device, chunk=1 << 24):
    import torch
    nd = len(shp)
    n = int(np.prod(shp))
    strides = [int(np.prod(shp[a + 1:])) for a in range(nd)]
    modes = []
    for _ in range(nmodes):
        lam = math.exp(rng.uniform(math.log(lam_lo), math.log(lam_hi)))
        v = rng.normal(size=nd)
        v /= np.linalg.norm(v) + 1e-300
        modes.append((lam, [float(2 * math.pi / lam * c) for c in v], float(rng.uniform(0, 2 * math.pi))))
    out = torch.empty(n, dtype=torch.float32, device=device)
    for c0 in range(0, n, chunk):
        idx = torch.arange(c0, min(n, c0 + chunk), device=device, dtype=torch.int64)
        coords = [((idx // strides[a]) % shp[a]).to(torch.float32) for a in range(nd)]
        acc = torch.zeros(idx.numel(), dtype=torch.float32, device=device)
        for lam, k, phi in modes:
            ph = torch.full_like(acc, phi)
            for a in range(nd):
                ph.add_(coords[a], alpha=k[a])
            acc.add_(torch.cos(ph), alpha=lam)
        out[c0:c0 + idx.numel()] = acc
    return out


def _quantile_gpu(t, q, gen, sample=1 << 20):
    import torch
    if t.numel() > sample:
        t = t[torch.randint(0, t.numel(), (sample,), device=t.device, generator=gen)]
    k = min(t.numel(), max(1, int(q * t.numel())))
    return torch.kthvalue(t.float().cpu(), k).values.item()


def _gen(device, seed):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def waves_gpu(dims, seed, device="cuda", nmodes=32, zero_frac=0.15, noise=1e-3):
    import torch
    rng, gen = np.random.default_rng(seed), _gen(device, seed)
    shp = _squeeze_shape(dims)
    f = _smooth_gpu(shp, rng, nmodes, 20.0, 900.0, device)
    lo, hi = f.min(), f.max()
    f.sub_(lo).div_(torch.clamp(hi - lo, min=1e-30))
    f.add_(torch.randn(f.numel(), device=device, generator=gen), alpha=noise).clamp_(0.0, 1.0)
    m = _smooth_gpu(shp, rng, 8, 300.0, max(300.0, 2.0 * max(shp)), device)
    f[m < _quantile_gpu(m, zero_frac, gen)] = 0.0
    return f


def peaks_gpu(dims, seed, device="cuda", zero_frac=0.97, bg_rows=0.05, bg_seg=0.40, bg_sigma=0.5):
    import torch
    rng, gen = np.random.default_rng(seed), _gen(device, seed)
    shp = _squeeze_shape(dims)
    s = _smooth_gpu(shp, rng, 32, 4.0, 40.0, device)
    thr = _quantile_gpu(s, zero_frac, gen)
    u = _smooth_gpu(shp, rng, 4, 200.0, max(200.0, 2.0 * max(shp)), device)
    u.sub_(u.min()).div_(torch.clamp(u.max() - u.min(), min=1e-30))
    h = torch.pow(10.0, 1.0 + 3.0 * u)
    f = torch.where(s > thr, (s - thr) / max(float(s.max()) - thr, 1e-30) * h, torch.zeros_like(s))
    rowlen = shp[-1]
    nrows = f.numel() // rowlen
    nbg = int(round(bg_rows * nrows)) if nrows > 1 else 0
    if nbg:
        seg = max(1, int(bg_seg * rowlen))
        rows = torch.randperm(nrows, device=device, generator=gen)[:nbg]
        start = torch.randint(0, rowlen - seg + 1, (nbg,), device=device, generator=gen)
        idx = (rows * rowlen + start).unsqueeze(1) + torch.arange(seg, device=device).unsqueeze(0)
        f.view(-1).index_add_(0, idx.reshape(-1), bg_sigma * torch.randn(idx.numel(), device=device, generator=gen))
    return f


def blobs_gpu(dims, seed, device="cuda", vmax=1e-3, zero_frac=0.70):
    import torch
    rng, gen = np.random.default_rng(seed), _gen(device, seed)
    shp = _squeeze_shape(dims)
    f = _smooth_gpu(shp, rng, 16, 20.0, 160.0, device)
    f.sub_(f.min())
    f.mul_(vmax / max(float(f.max()), 1e-30))
    f[f < _quantile_gpu(f, zero_frac, gen)] = 0.0
    return f


GPU_GENERATORS = {"waves": waves_gpu, "peaks": peaks_gpu, "blobs": blobs_gpu}


def generate_gpu(kind, dims, seed, device="cuda"):
    """float32 cuda tensor of x*y*z values (flat, x fastest)."""
    return GPU_GENERATORS[kind](dims, seed, device).contiguous()
