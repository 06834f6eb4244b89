"""Pins the CPU oracle (oracle/fz_oracle.c) against closed forms, known answers and brute force.

The reference publishes one known answer for this path (README.md:34-38: CESM 3600x1800 at
eb 1e-3 gives 3,105,556 bytes, CR 8.346332); its dataset is not available offline, so that
test is the size arithmetic of fz.cu:208 on its numbers.  Everything else is pinned by the
mathematics of each stage, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails at least one check.
"""
import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------- sizes (fz.cu:157-208)

def test_readme_known_answer_size_arithmetic():
    # CESM-ATM CLDHGH 3600x1800 f32: len = 6,480,000 floats, W = ceil(2 len / 512)
    n = 3600 * 1800
    W = O.flag_words(n)
    assert W == 25313
    assert O.tiles(n) == 3164 and n - 2048 * O.tiles(n) == 128
    K = 187769  # nonzero blocks implied by README.md:36
    csize = 4 * W + 16 * K
    assert csize == 3105556
    assert "%f" % (np.float32(4 * n) / np.float32(csize)) == "8.346332"


@pytest.mark.parametrize("n", [1, 255, 256, 257, 2047, 2048, 2049, 6480000, 15209600, 25000000])
def test_flag_words_formula(n):
    # dataChunkSize = ceil(2 len / (16 * 32)), fz.cu:177-178
    assert O.flag_words(n) == -(-2 * n // 512)
    assert O.tiles(n) == n // 2048


# ---------------------------------------------------------------- bitshuffle (fz.cu:61-111)

def _words(ntiles, rng, density=0.5):
    w = rng.integers(0, 2**32, size=1024 * ntiles, dtype=np.uint64).astype(np.uint32)
    keep = rng.random(w.size) < density
    return np.where(keep, w, 0).astype(np.uint32)


def test_bitshuffle_single_bit_lands_at_transposed_position():
    # bit b of word q (q = 32k + j) -> bit j of word 32b + k
    for q in (0, 1, 31, 32, 33, 500, 1023):
        for b in (0, 1, 7, 16, 31):
            w = np.zeros(1024, np.uint32)
            w[q] = np.uint32(1) << np.uint32(b)
            out, bf, F = O.bitshuffle_flags(w, 1)
            k, j = divmod(q, 32)
            expect = np.zeros(1024, np.uint32)
            expect[32 * b + k] = np.uint32(1) << np.uint32(j)
            assert np.array_equal(out, expect), (q, b)
            blk = (32 * b + k) // 4
            assert F.sum() == 1 and F[blk] == 1
            assert bf[blk // 32] == np.uint32(1) << np.uint32(blk % 32)


def test_bitshuffle_all_ones_and_zeros():
    out, bf, F = O.bitshuffle_flags(np.full(1024, 0xFFFFFFFF, np.uint32), 1)
    assert (out == 0xFFFFFFFF).all() and (F == 1).all() and (bf == 0xFFFFFFFF).all()
    out, bf, F = O.bitshuffle_flags(np.zeros(2048, np.uint32), 2)
    assert (out == 0).all() and (F == 0).all() and (bf == 0).all()


def test_bitshuffle_is_a_bit_permutation_of_order_three():
    rng = np.random.default_rng(7)
    w = _words(3, rng)
    o1, _, _ = O.bitshuffle_flags(w, 3)
    o2, _, _ = O.bitshuffle_flags(o1, 3)
    o3, _, _ = O.bitshuffle_flags(o2, 3)
    pc = lambda a: int(np.unpackbits(a.view(np.uint8)).sum())
    assert pc(o1) == pc(w)
    assert not np.array_equal(o1, w)
    assert np.array_equal(o3, w)  # (k, j, b) -> (b, k, j) is a 3-cycle
    assert np.array_equal(O.unshuffle(o1, 3), w)
    assert np.array_equal(o2, O.unshuffle(w, 3))


def test_flags_are_block_nonzero_and_lsb_first_packbits():
    rng = np.random.default_rng(8)
    w = _words(4, rng, density=0.01)
    out, bf, F = O.bitshuffle_flags(w, 4)
    blocks = out.reshape(-1, 4)
    assert np.array_equal(F, (blocks != 0).any(axis=1).astype(np.uint32))
    packed = np.packbits(F.astype(np.uint8), bitorder="little").view(np.uint32)
    assert np.array_equal(bf, packed)


# ---------------------------------------------------------------- scan + encode

def test_exclusive_sum_matches_cumsum():
    rng = np.random.default_rng(9)
    for n in (1, 5, 4096, 4097, 200000):
        a = rng.integers(0, 2, size=n, dtype=np.uint32)
        ex = O.exclusive_sum(a)
        ref = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(np.uint32)
        assert np.array_equal(ex, ref)
        assert np.array_equal(O.exclusive_sum(a, nthreads=4), ref)


def test_encode_is_stream_compaction():
    rng = np.random.default_rng(10)
    nb = 32 * 40
    data = rng.integers(0, 256, size=16 * nb, dtype=np.uint8)
    F = (rng.random(nb) < 0.3).astype(np.uint32)
    P = np.concatenate([[0], np.cumsum(F)]).astype(np.uint32)
    out = O.encode(data, P, nb)
    assert out.size == 16 * int(F.sum())
    assert np.array_equal(out.reshape(-1, 16), data.reshape(-1, 16)[F.astype(bool)])


# ---------------------------------------------------------------- eb parameters + prequant

def test_eb_params_narrow_double_to_float():
    for rng_, eb in [(1.0, 1e-3), (9857.9, 1e-3), (1e-3, 1e-4), (3.0, 0.1)]:
        eb_abs, r, e = O.eb_params(np.float32(rng_), eb)
        assert eb_abs == eb * float(np.float32(rng_))
        assert r == np.float32(1.0 / (2.0 * eb_abs))
        assert e == np.float32(2.0 * eb_abs)


def test_prequant_rounds_half_away_from_zero():
    # 1D, integer ramp with ebx2_r = 1: s = roundf(x); ties go away from zero
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.49999997, -0.49999997, 3.0], np.float32)
    codes, _ = O.dualquant(x, (x.size, 1, 1), np.float32(1.0))
    s = np.array([1, 2, 3, -1, -2, -3, 0, 0, 3], np.int64)
    d = np.diff(np.concatenate([[0], s]))
    expect = (np.abs(d) | ((d < 0).astype(np.int64) << 15)).astype(np.uint16)
    assert np.array_equal(codes, expect)


# ---------------------------------------------------------------- Lorenzo closed forms

def _code(d):
    d = np.asarray(d, np.int64)
    return (np.minimum(np.abs(d), 65535) | ((d < 0).astype(np.int64) << 15)).astype(np.uint16)


def test_lorenzo_1d_closed_form_and_literal_sign_loss():
    n = 1000
    s = (3 * np.arange(n) - 700).astype(np.int64)  # slope 3, negative start
    x = s.astype(np.float32)
    d = np.where(np.arange(n) % 256 == 0, s, 3)
    fixed, _ = O.dualquant(x, (n, 1, 1), np.float32(1.0), O.MODE_FIXED)
    assert np.array_equal(fixed, _code(d))
    # descending ramp: deltas are -3; LITERAL keeps the sign only at i % 4 == 0 (D1)
    s2 = -s
    lit, _ = O.dualquant(s2.astype(np.float32), (n, 1, 1), np.float32(1.0), O.MODE_LITERAL)
    d2 = np.where(np.arange(n) % 256 == 0, s2, -3)
    exp = _code(d2)
    nosign = (np.arange(n) % 4) != 0
    exp[nosign] &= 0x7FFF
    assert np.array_equal(lit, exp)


def test_lorenzo_2d_closed_form_on_a_plane():
    X, Y = 37, 35
    a, b, c = 5, -7, 11
    yy, xx = np.mgrid[0:Y, 0:X]
    s = a * xx + b * yy + c
    codes, _ = O.dualquant(s.astype(np.float32), (X, Y, 1), np.float32(1.0))
    bx, by = xx % 16 != 0, yy % 16 != 0
    d = np.where(bx & by, 0, np.where(bx, a, np.where(by, b, s)))
    assert np.array_equal(codes.reshape(Y, X), _code(d))


def test_lorenzo_3d_closed_form_on_a_hyperplane():
    X, Y, Z = 19, 10, 17
    a, b, c, e = 2, -3, 4, 9
    zz, yy, xx = np.mgrid[0:Z, 0:Y, 0:X]
    s = a * xx + b * yy + c * zz + e
    codes, _ = O.dualquant(s.astype(np.float32), (X, Y, Z), np.float32(1.0))
    bx, by, bz = xx % 8 != 0, yy % 8 != 0, zz % 8 != 0
    nb = bx.astype(int) + by.astype(int) + bz.astype(int)
    # a plane has zero 2nd/3rd differences: inside a block d = 0; on one face the 1D/2D
    # difference along the cut axes remains
    d = np.zeros_like(s)
    d = np.where(nb == 0, s, d)
    d = np.where((nb == 1) & bx, a, d)
    d = np.where((nb == 1) & by, b, d)
    d = np.where((nb == 1) & bz, c, d)
    assert np.array_equal(codes.reshape(Z, Y, X), _code(d))


def _naive_codes(x, dims, r):
    """Per-element Lorenzo with the reference's float op order (lorenzo_var.cuh:140-142, 277-283)."""
    X, Y, Z = dims
    f = x.reshape(Z, Y, X).astype(np.float32)
    t = (f * np.float32(r)).astype(np.float32)
    s = np.trunc(t + np.copysign(np.float32(0.49999997), t)).astype(np.float32)
    nd = O.ndim(X, Y, Z)
    out = np.zeros((Z, Y, X), np.uint16)
    z32 = np.float32(0)
    for zi in range(Z):
        for yi in range(Y):
            for xi in range(X):
                if nd == 1:
                    left = s[0, 0, xi - 1] if xi % 256 else z32
                    d = np.float32(s[0, 0, xi] - left)
                elif nd == 2:
                    bx, by = xi % 16 != 0, yi % 16 != 0
                    N = s[0, yi - 1, xi] if by else z32
                    NW = s[0, yi - 1, xi - 1] if (by and bx) else z32
                    W = s[0, yi, xi - 1] if bx else z32
                    c1 = np.float32(s[0, yi, xi] - N)
                    w1 = np.float32(W - NW)
                    d = np.float32(c1 - w1) if bx else c1
                else:
                    bx, by, bz = xi % 8 != 0, yi % 8 != 0, zi % 8 != 0
                    g = lambda dz, dy, dx, ok: s[zi - dz, yi - dy, xi - dx] if ok else z32
                    p = g(1, 1, 1, bz and by and bx)
                    p = np.float32(p - g(0, 1, 1, by and bx))
                    p = np.float32(p - g(1, 0, 1, bz and bx))
                    p = np.float32(p - g(1, 1, 0, bz and by))
                    p = np.float32(p + g(0, 0, 1, bx))
                    p = np.float32(p + g(0, 1, 0, by))
                    p = np.float32(p + g(1, 0, 0, bz))
                    d = np.float32(s[zi, yi, xi] - p)
                out[zi, yi, xi] = _code(int(d))
    return out.ravel()


@pytest.mark.parametrize("dims", [(300, 1, 1), (16, 16, 1), (17, 33, 1), (33, 17, 1), (8, 8, 8), (9, 9, 9),
                                  (20, 1, 9), (1, 20, 1), (1, 1, 1)])
def test_lorenzo_brute_force_small_shapes(dims):
    rng = np.random.default_rng(sum(dims))
    n = dims[0] * dims[1] * dims[2]
    x = (rng.normal(size=n) * 40).astype(np.float32)
    x[rng.random(n) < 0.2] = 0
    r = np.float32(1.0 / (2.0 * 0.37))
    codes, _ = O.dualquant(x, dims, r)
    assert np.array_equal(codes, _naive_codes(x, dims, r))


@pytest.mark.parametrize("dims", [(1000, 1, 1), (37, 35, 1), (19, 10, 17), (64, 1, 20)])
def test_reconstruct_inverts_lorenzo_exactly(dims):
    rng = np.random.default_rng(3)
    n = dims[0] * dims[1] * dims[2]
    s = rng.integers(-3000, 3000, size=n).astype(np.float32)
    codes, ovf = O.dualquant(s, dims, np.float32(1.0))
    assert ovf == 0
    back = O.reconstruct(codes, dims, np.float32(1.0))
    assert np.array_equal(back, s)


# ---------------------------------------------------------------- whole pipeline

def _field(dims, seed, zero=0.3):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    x = np.cumsum(rng.normal(size=n)).astype(np.float32)
    x[rng.random(n) < zero] = 0
    return x


@pytest.mark.parametrize("dims", [(5000, 1, 1), (100, 77, 1), (33, 20, 17), (4096, 1, 1), (16, 16, 1), (64, 64, 1)])
def test_compress_decompress_roundtrip_and_invariants(dims):
    x = _field(dims, 11)
    n = x.size
    r = O.compress(x, dims, 1e-3, O.MODE_FIXED, debug=True)
    assert r["status"] == 0
    info = r["info"]
    W, T = O.flag_words(n), O.tiles(n)
    B = r["payload"][: 4 * W].view(np.uint32)
    # popcount(B) = K, size formula fz.cu:208, no flag past the last full tile
    assert int(np.unpackbits(B.view(np.uint8)).sum()) == info["K"]
    assert info["payload_bytes"] == 4 * W + 16 * info["K"]
    assert (B[8 * T:] == 0).all()
    assert r["tail"].size == n - 2048 * T
    rc, xr = O.decompress(r["payload"], r["tail"], n, dims, info["ebx2"])
    assert rc == 0
    tol = info["eb_abs"] + 4 * 2.0**-24 * float(np.abs(x).max())
    assert float(np.abs(xr - x).max()) <= tol


def test_fixed_and_literal_payloads_agree_for_2d_3d():
    for dims in [(100, 77, 1), (33, 20, 17)]:
        x = _field(dims, 12)
        a = O.compress(x, dims, 1e-3, O.MODE_FIXED)
        b = O.compress(x, dims, 1e-3, O.MODE_LITERAL)
        assert np.array_equal(a["payload"], b["payload"])
        assert b["tail"].size == 0


def test_openmp_path_is_byte_identical():
    for dims in [(300000, 1, 1), (700, 500, 1), (60, 50, 40)]:
        x = _field(dims, 13)
        a = O.compress(x, dims, 1e-4, nthreads=1, debug=True)
        b = O.compress(x, dims, 1e-4, nthreads=4, debug=True)
        for k in ("payload", "tail", "codes", "byteflags", "presum"):
            assert np.array_equal(a[k], b[k]), k


def test_error_statuses():
    assert O.compress(np.full(5000, 3.0, np.float32), (5000, 1, 1), 1e-3)["status"] == 4   # D5
    x = _field((5000, 1, 1), 1)
    x[17] = np.nan
    assert O.compress(x, (5000, 1, 1), 1e-3)["status"] == 5
    assert O.compress(_field((100, 1, 1), 1), (101, 1, 1), 1e-3)["status"] == 1           # len < xyz
    assert O.compress(_field((100, 1, 1), 1), (100, 1, 1), 0.0)["status"] == 1


def test_overflow_is_counted():
    x = np.zeros(4096, np.float32)
    x[100] = 1.0
    r = O.compress(x, (4096, 1, 1), 1e-6)  # step of 1 at eb 1e-6 -> |d| = 500000
    assert r["status"] == 0 and r["info"]["overflow"] == 2


def test_len_longer_than_dims_gives_zero_codes():
    x = _field((3000, 1, 1), 5)
    r = O.compress(x, (1000, 2, 1), 1e-3, debug=True)
    assert r["status"] == 0
    assert (r["codes"][2000:] == 0).all()
