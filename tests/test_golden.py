"""The committed golden fixtures (tests/golden/, written by make_golden.py) against the oracle
and the input generators on the CPU: every frozen intermediate (codes, byte flags, their scan,
payload, tail, decompressed floats, side channel, metrics) is reproduced bit for bit, so a
change of the oracle or of the seeded inputs shows up here before any GPU comparison."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _npz(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unhex(v):
    return float.fromhex(v) if isinstance(v, str) else v


def same_record(info, rec):
    for k, v in rec.items():
        a, b = info[k], unhex(v)
        assert (math.isnan(a) and math.isnan(b)) or a == b, (k, a, b)


SMALL = _json("small_cases.json")
OUTL = _json("outlier_cases.json")


def small_field(n, seed):
    rng = np.random.default_rng(seed)
    return (np.cumsum(rng.normal(size=n)) + 0.5 * np.sin(np.arange(n) / 40.0)).astype(np.float32)


@pytest.mark.parametrize("name", sorted(SMALL))
def test_small_case_reproduced_by_oracle(name):
    m, z = SMALL[name], _npz("small_cases.npz")
    dims, n, eb, mode = tuple(m["dims"]), m["len"], float.fromhex(m["eb_rel"]), m["mode"]
    x = z[f"{name}/x"]
    assert np.array_equal(x.view(np.uint32), small_field(n, m["seed"]).view(np.uint32))
    r = O.compress(x, dims, eb, mode, debug=True)
    same_record(r["info"], m["record"])
    W = r["info"]["W"]
    assert np.array_equal(r["codes"], z[f"{name}/codes"])
    assert np.array_equal(r["byteflags"][: 32 * W], z[f"{name}/byteflags"])
    assert np.array_equal(r["presum"][: 32 * W + 1], z[f"{name}/presum"])
    assert np.array_equal(r["payload"], z[f"{name}/payload"])
    assert np.array_equal(r["tail"], z[f"{name}/tail"])
    # the frozen flags and scan agree with the frozen payload: B = packbits(F), K = P[N]
    F = z[f"{name}/byteflags"]
    B = np.packbits(F.astype(np.uint8), bitorder="little").view(np.uint32)
    assert np.array_equal(B, z[f"{name}/payload"][: 4 * W].view(np.uint32))
    assert z[f"{name}/presum"][-1] == m["record"]["K"] == int(F.sum())
    rc, xhat = O.decompress(z[f"{name}/payload"], z[f"{name}/tail"] if mode == O.MODE_FIXED else None, n, dims,
                            unhex(m["record"]["ebx2"]))
    assert rc == 0 and np.array_equal(xhat.view(np.uint32), z[f"{name}/xhat"].view(np.uint32))


@pytest.mark.parametrize("name", sorted(OUTL))
def test_outlier_case_reproduced_by_oracle(name):
    m, z = OUTL[name], _npz("outlier_cases.npz")
    dims, n, eb = tuple(m["dims"]), m["len"], float.fromhex(m["eb_rel"])
    r = O.compress_outlier(z[f"{name}/x"], dims, eb)
    same_record(r["info"], m["record"])
    for k in ("payload", "tail", "side"):
        assert np.array_equal(r[k], z[f"{name}/{k}"]), k
    rc, xhat = O.decompress_outlier(z[f"{name}/payload"], z[f"{name}/tail"], z[f"{name}/side"], n, dims,
                                    unhex(m["record"]["ebx2"]), unhex(m["record"]["vmin"]))
    assert rc == 0 and np.array_equal(xhat.view(np.uint32), z[f"{name}/xhat"].view(np.uint32))
    # every non-finite input comes back bit for bit
    x = z[f"{name}/x"]
    bad = ~np.isfinite(x)
    assert bad.any() and np.array_equal(xhat[bad].view(np.uint32), x[bad].view(np.uint32))


def test_verify_cases_reproduced_by_oracle():
    z = _npz("small_cases.npz")
    for name, want in _json("verify_cases.json").items():
        if name == "constant_identical":
            x = o = np.full(3000, 2.5, np.float32)
        else:
            xyz = int(np.prod(SMALL[name]["dims"]))
            x, o = z[f"{name}/xhat"][:xyz], z[f"{name}/x"][:xyz]
        got = O.verify(x, o)
        for grp, vals in want.items():
            if grp == "len":
                assert got["len"] == vals
                continue
            for k, v in vals.items():
                a, b = got[grp][k], unhex(v)
                assert (math.isnan(a) and math.isnan(b)) or a == b, (name, grp, k, a, b)


@pytest.mark.parametrize("i", range(4))
def test_benchmark_fields_and_their_archives(i):
    from fzgpu import datagen as D
    c = _json("configs.json")[i]
    x, dims, _ = D.config_field(c["workload"])
    x = x.ravel()
    assert tuple(dims) == tuple(c["dims"]) and _sha(x) == c["input_sha256"]
    r = O.compress(x, dims, float.fromhex(c["eb_rel"]), O.MODE_FIXED, nthreads=8)
    same_record(r["info"], c["record"])
    assert _sha(r["payload"]) == c["payload_sha256"] and _sha(r["tail"]) == c["tail_sha256"]
    rc, xhat = O.decompress(r["payload"], r["tail"], x.size, dims, r["info"]["ebx2"], nthreads=8)
    assert rc == 0 and _sha(xhat) == c["xhat_sha256"]
