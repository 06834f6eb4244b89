"""The HIP path against the committed golden fixtures (tests/golden/): the fused compress, the
plain per-stage pipeline, the per-stage Lorenzo codes, decompress, the outlier mode and the
quality metrics on the frozen small cases byte for byte; the three benchmark fields by the
SHA-256 of their payload, tail and decompressed floats."""
import hashlib
import json
import math
import os
import sys

import numpy as np
import pytest

import fzgpu

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _npz(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _np(t):
    return t.cpu().numpy()


def unhex(v):
    return float.fromhex(v) if isinstance(v, str) else v


# the result record of fz_compress_f32 under the oracle's names
KEYS = {"status": "status", "ndim": "ndim", "vmin": "vmin", "vmax": "vmax", "ebx2_r": "ebx2_r", "ebx2": "ebx2",
        "eb_abs": "eb_abs", "nblocks": "K", "payload_bytes": "payload_bytes", "tail_len": "tail_len",
        "overflow": "overflow"}


def same_record(r, rec, literal=False):
    for k_gpu, k_gold in KEYS.items():
        a, b = r[k_gpu], unhex(rec[k_gold])
        if literal and k_gpu == "tail_len":  # LITERAL mode returns no tail (the reference drops it)
            b = 0
        assert (math.isnan(a) and math.isnan(b)) or a == b, (k_gpu, a, b)


SMALL = _json("small_cases.json")
OUTL = _json("outlier_cases.json")


@pytest.mark.parametrize("name", sorted(SMALL))
def test_small_case_on_gpu(name):
    m, z = SMALL[name], _npz("small_cases.npz")
    dims, n, eb, mode = tuple(m["dims"]), m["len"], float.fromhex(m["eb_rel"]), m["mode"]
    say(f"case golden {name} dims={dims} len={n}")
    d = _dev(z[f"{name}/x"])
    plan = fzgpu.Plan(*dims, length=n, mode=mode)
    for fused in (True, False):
        payload, tail, r = plan.compress(d, eb, fused=fused)
        same_record(r, m["record"], literal=mode == fzgpu.MODE_LITERAL)
        assert np.array_equal(_np(payload), z[f"{name}/payload"]), fused
        if mode == fzgpu.MODE_FIXED:
            assert np.array_equal(_np(tail).view(np.uint16), z[f"{name}/tail"]), fused
    codes, _ = fzgpu.lorenzo_construct(d, dims, unhex(m["record"]["eb_abs"]), mode=mode, length=n)
    assert np.array_equal(_np(codes)[:n].view(np.uint16), z[f"{name}/codes"])
    t = _dev(z[f"{name}/tail"].view(np.int16)) if mode == fzgpu.MODE_FIXED else None
    out = plan.decompress(_dev(z[f"{name}/payload"]), t, unhex(m["record"]["ebx2"]))
    assert np.array_equal(_np(out).view(np.uint32), z[f"{name}/xhat"].view(np.uint32))


@pytest.mark.parametrize("name", sorted(OUTL))
def test_outlier_case_on_gpu(name):
    m, z = OUTL[name], _npz("outlier_cases.npz")
    dims, eb = tuple(m["dims"]), float.fromhex(m["eb_rel"])
    say(f"case golden outlier {name} dims={dims}")
    plan = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER)
    payload, tail, r, side = plan.compress_outlier(_dev(z[f"{name}/x"]), eb)
    same_record(r, m["record"])
    assert np.array_equal(_np(payload), z[f"{name}/payload"])
    assert np.array_equal(_np(tail).view(np.uint16), z[f"{name}/tail"])
    assert np.array_equal(_np(side), z[f"{name}/side"])
    out = plan.decompress_outlier(_dev(z[f"{name}/payload"]), _dev(z[f"{name}/tail"].view(np.int16)),
                                  _dev(z[f"{name}/side"]), unhex(m["record"]["ebx2"]), unhex(m["record"]["vmin"]))
    assert np.array_equal(_np(out).view(np.uint32), z[f"{name}/xhat"].view(np.uint32))


# extrema, ranges, the largest error, its index and the ratios exactly; sums-derived values to
# a relative 1e-9 and PSNR to 1e-9 dB (the device adds in another order)
EXACT = {("odata", "min"), ("odata", "max"), ("odata", "rng"), ("xdata", "min"), ("xdata", "max"),
         ("xdata", "rng"), ("max_err", "abs"), ("max_err", "idx"), ("max_err", "rel"), ("max_err", "pwrrel")}


def test_verify_cases_on_gpu():
    z = _npz("small_cases.npz")
    for name, want in _json("verify_cases.json").items():
        say(f"case golden verify {name}")
        if name == "constant_identical":
            x = o = np.full(3000, 2.5, np.float32)
        else:
            xyz = int(np.prod(SMALL[name]["dims"]))
            x, o = z[f"{name}/xhat"][:xyz], z[f"{name}/x"][:xyz]
        got = fzgpu.verify(_dev(x), _dev(o))
        assert got["len"] == want["len"]
        for grp in ("odata", "xdata", "reduced", "max_err"):
            for k, v in want[grp].items():
                a, b = float(got[grp][k]), float(unhex(v))
                if math.isnan(b) or math.isinf(b) or (grp, k) in EXACT:
                    assert (math.isnan(a) and math.isnan(b)) or a == b, (name, grp, k, a, b)
                elif k == "PSNR":
                    assert abs(a - b) <= 1e-9, (name, k, a, b)
                else:
                    assert abs(a - b) <= 1e-9 * abs(b), (name, grp, k, a, b)


@pytest.mark.parametrize("i", range(4))
def test_benchmark_field_archive_hashes(i):
    from fzgpu import datagen as D
    c = _json("configs.json")[i]
    x, dims, _ = D.config_field(c["workload"])
    x = x.ravel()
    assert _sha(x) == c["input_sha256"]
    eb = float.fromhex(c["eb_rel"])
    say(f"case golden config {c['workload']} eb={eb}")
    plan = fzgpu.Plan(*dims)
    payload, tail, r = plan.compress(_dev(x), eb)
    same_record(r, c["record"])
    assert _sha(_np(payload)) == c["payload_sha256"]
    assert _sha(_np(tail).view(np.uint16)) == c["tail_sha256"]
    out = plan.decompress(payload, tail, r["ebx2"])
    assert _sha(_np(out)) == c["xhat_sha256"]
