"""The oracle's OUTLIER mode (the side channel for |d| >= 32768, non-finite values, values too
large to quantise exactly, and constant fields) pinned by round trips, by the FIXED-mode
codes where the reference's float arithmetic is exact, and by a numpy restatement of the
block-local integer Lorenzo."""
import numpy as np
import pytest

import oracle as O


def _round_half_away(t):
    t = t.astype(np.float64)
    return (np.sign(t) * np.floor(np.abs(t) + 0.5)).astype(np.int64)


def _deltas_numpy(q, dims):
    """Block-local Lorenzo deltas of the integer field q (x fastest), by array shifts."""
    x, y, z = dims
    if y == 1 and z == 1:
        d = q.copy()
        left = np.zeros_like(q)
        left[1:] = q[:-1]
        left[np.arange(q.size) % 256 == 0] = 0
        return d - left
    if z == 1:
        a = q.reshape(y, x)
        ix, iy = np.meshgrid(np.arange(x), np.arange(y))
        bx, by = (ix % 16) != 0, (iy % 16) != 0
        w = np.zeros_like(a); w[:, 1:] = a[:, :-1]
        n = np.zeros_like(a); n[1:, :] = a[:-1, :]
        nw = np.zeros_like(a); nw[1:, 1:] = a[:-1, :-1]
        return (a - np.where(bx, w, 0) - np.where(by, n, 0) + np.where(bx & by, nw, 0)).ravel()
    a = q.reshape(z, y, x)
    iz, iy, ix = np.meshgrid(np.arange(z), np.arange(y), np.arange(x), indexing="ij")
    bx, by, bz = (ix % 8) != 0, (iy % 8) != 0, (iz % 8) != 0

    def sh(dz, dy, dx):
        o = np.zeros_like(a)
        o[dz:, dy:, dx:] = a[: z - dz, : y - dy, : x - dx]
        return o
    d = (a - np.where(bx, sh(0, 0, 1), 0) - np.where(by, sh(0, 1, 0), 0) - np.where(bz, sh(1, 0, 0), 0)
         + np.where(bx & by, sh(0, 1, 1), 0) + np.where(bx & bz, sh(1, 0, 1), 0) + np.where(by & bz, sh(1, 1, 0), 0)
         - np.where(bx & by & bz, sh(1, 1, 1), 0))
    return d.ravel()


def _field(dims, seed, spikes=0, nans=0, infs=0, scale=1.0):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    x = (np.cumsum(rng.normal(size=n)) * scale).astype(np.float32)
    if spikes:
        x[rng.integers(0, n, spikes)] = np.float32(1e7 * scale)
    if nans:
        x[rng.integers(0, n, nans)] = np.nan
    if infs:
        x[rng.integers(0, n, infs)] = np.float32(np.inf)
        x[rng.integers(0, n, infs)] = np.float32(-np.inf)
    return x


def _check_round_trip(x, dims, eb):
    r = O.compress_outlier(x, dims, eb)
    assert r["status"] == 0
    info = r["info"]
    rc, out = O.decompress_outlier(r["payload"], r["tail"], r["side"], x.size, dims, info["ebx2"], info["vmin"])
    assert rc == 0
    fin = np.isfinite(x)
    bound = info["eb_abs"] + 4 * 2.0 ** -24 * float(np.abs(x[fin]).max(initial=0.0))
    assert np.abs(out[fin].astype(np.float64) - x[fin]).max(initial=0.0) <= bound
    # every verbatim value comes back bit for bit
    assert np.array_equal(out[~fin].view(np.uint32), x[~fin].view(np.uint32))
    return r, out


SHAPES = [(5000, 1, 1), (4097, 1, 1), (100, 77, 1), (33, 20, 17), (64, 64, 64), (9, 1, 700)]


@pytest.mark.parametrize("dims", SHAPES)
def test_side_channel_lists_match_a_numpy_restatement(dims):
    x = _field(dims, 1, spikes=6, nans=3, infs=2)
    eb = 1e-6
    r, out = _check_round_trip(x, dims, eb)
    info = r["info"]
    q_t = (x * np.float32(info["ebx2_r"])).astype(np.float32)
    raw = ~np.isfinite(x) | ~(np.abs(q_t) < 2.0 ** 24)
    q = np.where(raw, 0, _round_half_away(np.where(raw, 0, q_t)))
    d = _deltas_numpy(q, dims)
    big = np.abs(d) >= 32768
    idx_d, val_d, idx_r, val_r = O.side_lists(r["side"])
    assert np.array_equal(idx_d, np.nonzero(big)[0]) and np.array_equal(val_d, d[big])
    assert np.array_equal(idx_r, np.nonzero(raw)[0]) and np.array_equal(val_r, x[raw].view(np.uint32))
    assert info["overflow"] == big.sum() and info["n_raw"] == raw.sum() and big.sum() > 0
    codes = O.compress_outlier(x, dims, eb, debug=True)["codes"]
    small = ~big
    exp = np.where(d < 0, (-d) | 0x8000, d)
    assert np.array_equal(codes[: d.size][small], exp[small]) and not codes[: d.size][big].any()


@pytest.mark.parametrize("dims", SHAPES)
def test_codes_equal_the_fixed_mode_where_the_reference_arithmetic_is_exact(dims):
    # no raw values, no overflow, |s| < 2^21: the reference's float deltas are exact integers
    x = _field(dims, 2)
    for eb in (1e-2, 1e-3):
        fixed = O.compress(x, dims, eb, O.MODE_FIXED)
        outl = O.compress_outlier(x, dims, eb)
        assert outl["info"]["overflow"] == 0 and outl["info"]["n_raw"] == 0
        assert np.array_equal(fixed["payload"], outl["payload"]) and np.array_equal(fixed["tail"], outl["tail"])
        assert outl["side"].size == 16


def test_constant_fields_and_all_non_finite():
    dims = (300, 40, 1)
    c = np.full(12000, 3.25, np.float32)
    r, out = _check_round_trip(c, dims, 1e-3)
    assert r["info"]["ebx2"] == 0.0 and np.array_equal(out, c) and r["side"].size == 16
    c[[5, 7000]] = np.nan
    r, out = _check_round_trip(c, dims, 1e-3)
    assert r["info"]["n_raw"] == 2 and np.isnan(out[[5, 7000]]).all() and (np.delete(out, [5, 7000]) == 3.25).all()
    a = np.full(5000, np.nan, np.float32)
    r, out = _check_round_trip(a, (5000, 1, 1), 1e-3)
    assert r["info"]["n_raw"] == 5000 and (r["info"]["vmin"], r["info"]["vmax"]) == (0.0, 0.0)


def test_values_too_large_to_quantise_exactly_are_kept_verbatim():
    # 1e9 + noise with a unit range: |x * ebx2_r| ~ 5e11 >= 2^24.  The FIXED mode loses the
    # error bound there (its integers are rounded floats); the OUTLIER mode keeps every value.
    dims = (4096, 1, 1)
    rng = np.random.default_rng(3)
    x = (np.float32(1e9) + rng.integers(0, 64, 4096).astype(np.float32) * np.float32(64)).astype(np.float32)
    fixed = O.compress(x, dims, 1e-3, O.MODE_FIXED)
    rc, xf = O.decompress(fixed["payload"], fixed["tail"], x.size, dims, fixed["info"]["ebx2"])
    assert np.abs(xf.astype(np.float64) - x).max() > 10 * fixed["info"]["eb_abs"]
    r, out = _check_round_trip(x, dims, 1e-3)
    assert r["info"]["n_raw"] == x.size and np.array_equal(out, x)


def test_capacity_and_corrupt_side_channels():
    dims = (100, 77, 1)
    x = _field(dims, 4, spikes=4, nans=3)
    full = O.compress_outlier(x, dims, 1e-6)
    need = full["info"]["side_bytes"]
    assert O.compress_outlier(x, dims, 1e-6, side_cap=need - 1)["status"] == 7
    assert O.compress_outlier(x, dims, 1e-6, side_cap=need)["status"] == 0
    info = full["info"]
    bad = full["side"].copy()
    bad[16:20] = np.frombuffer(np.uint32(10 ** 6).tobytes(), np.uint8)  # index beyond the field
    assert O.decompress_outlier(full["payload"], full["tail"], bad, x.size, dims, info["ebx2"], info["vmin"])[0] == 6
    assert O.decompress_outlier(full["payload"], full["tail"], full["side"][:-4], x.size, dims, info["ebx2"],
                                info["vmin"])[0] == 6
