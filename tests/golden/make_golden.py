"""Writes the golden fixtures of this directory from the oracle (run on the CPU, from the repo
root: `python tests/golden/make_golden.py`).  The reference holds no test vectors of its own
(its tests print without expected values, and its one known answer, README.md:34-38, needs a
dataset that is not here), so the fixtures are the oracle's outputs on seeded inputs, frozen:

  small_cases.npz / small_cases.json
      small fields (1D / 2D / 3D, ragged, no full tile, len > x*y*z) in FIXED and LITERAL
      mode: the input, the codes, the byte flags and their exclusive scan, the payload
      (bitflags || nonzero blocks), the tail codes, the decompressed floats, and the result
      record (fz.cu:113-209; lorenzo_var.cuh:47-295).
  outlier_cases.npz / outlier_cases.json
      the outlier mode on fields with spikes, NaN and +-Inf: payload, tail, side channel and
      the decompressed floats.
  verify_cases.json
      verify_data (verify.hh:25-83) of decompressed small cases against their inputs.
  configs.json
      the three benchmark fields (datagen.config_field) at their error bounds (and the 3D one
      at 1e-4): SHA-256 of the input, payload, tail and decompressed floats, and the record.

Doubles are stored with float.hex so that they round-trip exactly.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "fz-gpu_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402

# name: (dims, length or 0, eb_rel, mode, seed)
SMALL = {
    "1d_5000_fixed": ((5000, 1, 1), 0, 1e-3, O.MODE_FIXED, 11),
    "1d_5000_literal": ((5000, 1, 1), 0, 1e-3, O.MODE_LITERAL, 11),
    "1d_2047_no_tile": ((2047, 1, 1), 0, 1e-3, O.MODE_FIXED, 12),
    "2d_100x77": ((100, 77, 1), 0, 1e-4, O.MODE_FIXED, 13),
    "2d_17x33_literal": ((17, 33, 1), 0, 1e-3, O.MODE_LITERAL, 14),
    "2d_1000x3_len6000": ((1000, 3, 1), 6000, 1e-3, O.MODE_FIXED, 15),
    "3d_33x20x17": ((33, 20, 17), 0, 1e-3, O.MODE_FIXED, 16),
    "3d_9x50x40": ((9, 50, 40), 0, 1e-4, O.MODE_FIXED, 17),
    "3d_24x20x18_literal": ((24, 20, 18), 0, 1e-3, O.MODE_LITERAL, 18),
}

# name: (dims, eb_rel, seed, spikes, nans, infs)
OUTLIER = {
    "1d_5000": ((5000, 1, 1), 1e-6, 21, 5, 4, 2),
    "2d_100x77": ((100, 77, 1), 1e-5, 22, 5, 4, 2),
    "3d_33x20x17": ((33, 20, 17), 1e-6, 23, 5, 4, 2),
}

CONFIG_RUNS = [("cesm-like-3600x1800", 1e-3), ("detector-39200x388", 1e-3), ("blobs3d-500x500x100", 1e-3),
               ("blobs3d-500x500x100", 1e-4)]

RECORD = ("status", "ndim", "vmin", "vmax", "ebx2_r", "ebx2", "eb_abs", "K", "payload_bytes", "tail_len", "overflow")


def small_field(n, seed):
    rng = np.random.default_rng(seed)
    return (np.cumsum(rng.normal(size=n)) + 0.5 * np.sin(np.arange(n) / 40.0)).astype(np.float32)


def outlier_field(n, seed, spikes, nans, infs):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.normal(size=n)).astype(np.float32)
    x[rng.integers(0, n, spikes)] = np.float32(1e7)
    x[rng.integers(0, n, nans)] = np.nan
    x[rng.integers(0, n, infs)] = np.float32(np.inf)
    x[rng.integers(0, n, infs)] = np.float32(-np.inf)
    return x


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def hexed(v):
    return float(v).hex() if isinstance(v, float) else int(v)


def record(info):
    return {k: hexed(info[k]) for k in RECORD}


def stats_json(s):
    return {grp: ({k: hexed(v) for k, v in s[grp].items()} if isinstance(s[grp], dict) else int(s[grp]))
            for grp in ("odata", "xdata", "reduced", "max_err", "len")}


def main():
    from fzgpu import datagen as D
    arrays, meta, verify_cases = {}, {}, {}
    for name, (dims, length, eb, mode, seed) in SMALL.items():
        n = length or dims[0] * dims[1] * dims[2]
        x = small_field(n, seed)
        r = O.compress(x, dims, eb, mode, debug=True)
        assert r["status"] == 0, name
        info = r["info"]
        rc, xhat = O.decompress(r["payload"], r["tail"] if mode == O.MODE_FIXED else None, n, dims, info["ebx2"])
        assert rc == 0, name
        W = info["W"]
        arrays.update({f"{name}/x": x, f"{name}/codes": r["codes"], f"{name}/byteflags": r["byteflags"][: 32 * W],
                       f"{name}/presum": r["presum"][: 32 * W + 1], f"{name}/payload": r["payload"],
                       f"{name}/tail": r["tail"], f"{name}/xhat": xhat})
        meta[name] = {"dims": list(dims), "len": n, "eb_rel": eb.hex(), "mode": mode, "seed": seed,
                      "record": record(info)}
        if name in ("2d_100x77", "3d_9x50x40"):
            verify_cases[name] = stats_json(O.verify(xhat[: dims[0] * dims[1] * dims[2]],
                                                     x[: dims[0] * dims[1] * dims[2]]))
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **arrays)
    with open(os.path.join(HERE, "small_cases.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)

    arrays, meta = {}, {}
    for name, (dims, eb, seed, spikes, nans, infs) in OUTLIER.items():
        n = dims[0] * dims[1] * dims[2]
        x = outlier_field(n, seed, spikes, nans, infs)
        r = O.compress_outlier(x, dims, eb)
        assert r["status"] == 0, name
        info = r["info"]
        rc, xhat = O.decompress_outlier(r["payload"], r["tail"], r["side"], n, dims, info["ebx2"], info["vmin"])
        assert rc == 0, name
        arrays.update({f"{name}/x": x, f"{name}/payload": r["payload"], f"{name}/tail": r["tail"],
                       f"{name}/side": r["side"], f"{name}/xhat": xhat})
        rec = record(info)
        rec.update(n_raw=int(info["n_raw"]), side_bytes=int(info["side_bytes"]))
        meta[name] = {"dims": list(dims), "len": n, "eb_rel": eb.hex(), "seed": seed, "record": rec}
    np.savez_compressed(os.path.join(HERE, "outlier_cases.npz"), **arrays)
    with open(os.path.join(HERE, "outlier_cases.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)

    const = np.full(3000, 2.5, np.float32)
    verify_cases["constant_identical"] = stats_json(O.verify(const, const))
    with open(os.path.join(HERE, "verify_cases.json"), "w") as f:
        json.dump(verify_cases, f, indent=1, sort_keys=True)

    configs = []
    for name, eb in CONFIG_RUNS:
        x, dims, _ = D.config_field(name)
        x = x.ravel()
        r = O.compress(x, dims, eb, O.MODE_FIXED, nthreads=8)
        assert r["status"] == 0, name
        info = r["info"]
        rc, xhat = O.decompress(r["payload"], r["tail"], x.size, dims, info["ebx2"], nthreads=8)
        assert rc == 0, name
        configs.append({"workload": name, "dims": list(dims), "eb_rel": eb.hex(), "input_sha256": sha(x),
                        "payload_sha256": sha(r["payload"]), "tail_sha256": sha(r["tail"]),
                        "xhat_sha256": sha(xhat), "record": record(info)})
        print(name, eb, info["payload_bytes"], flush=True)
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(configs, f, indent=1)


if __name__ == "__main__":
    main()
