"""The oracle's verify_data restatement (verify.hh:25-83) pinned against closed forms, numpy's
own statistics routines and hand-worked cases of the reference's scan rules."""
import math

import numpy as np
import pytest

import oracle


def test_shifted_copy_closed_forms():
    # o = 0..n-1, x = o + 1/4: every error is 1/4 exactly
    n = 1000
    o = np.arange(n, dtype=np.float32)
    x = o + np.float32(0.25)
    s = oracle.verify(x, o)
    assert s["len"] == n
    assert s["odata"] == {"min": 0.0, "max": n - 1.0, "rng": n - 1.0, "std": pytest.approx(math.sqrt((n * n - 1) / 12), rel=1e-14)}
    assert s["xdata"]["std"] == pytest.approx(math.sqrt((n * n - 1) / 12), rel=1e-14)
    assert s["reduced"]["MSE"] == 0.0625
    assert s["reduced"]["NRMSE"] == pytest.approx(0.25 / (n - 1), rel=1e-15)
    assert s["reduced"]["PSNR"] == pytest.approx(20 * math.log10(n - 1) - 10 * math.log10(0.0625), rel=1e-14)
    assert s["reduced"]["coeff"] == pytest.approx(1.0, rel=1e-14)
    # every |x-o| ties: the strict scan keeps index 0
    assert (s["max_err"]["abs"], s["max_err"]["idx"]) == (0.25, 0)
    assert s["max_err"]["rel"] == pytest.approx(0.25 / (n - 1), rel=1e-15)
    # o = 0 is skipped; the largest |x-o|/|o| is at o = 1
    assert s["max_err"]["pwrrel"] == 0.25


def test_xdata_max_follows_the_reference_scan_of_odata():
    # verify.hh:44 compares odata against the xdata maximum: it is max(x[0], max odata)
    o = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    x = o + np.float32(0.5)
    s = oracle.verify(x, o)
    assert s["xdata"]["max"] == 4.0 and float(x.max()) == 4.5
    assert s["xdata"]["min"] == 1.5
    assert s["xdata"]["rng"] == 2.5
    x2 = x.copy()
    x2[0] = 10.0
    assert oracle.verify(x2, o)["xdata"]["max"] == 10.0


def test_first_index_of_the_largest_error_and_anticorrelation():
    o = np.linspace(-1, 1, 101).astype(np.float32)
    x = -o
    x[7] += np.float32(0.5)
    s = oracle.verify(x, o)
    err = np.abs(x - o)
    assert s["max_err"]["idx"] == int(np.argmax(err))  # numpy's argmax keeps the first maximum
    assert s["max_err"]["abs"] == float(err.max())
    y = -o
    s = oracle.verify(y, o)
    assert s["reduced"]["coeff"] == pytest.approx(-1.0, rel=1e-13)
    # |x-o| = 2|o| peaks at both ends; the first end wins
    assert s["max_err"]["idx"] == 0
    assert s["max_err"]["pwrrel"] == 2.0


def test_zeros_of_odata_are_skipped_by_the_pointwise_ratio():
    o = np.array([0.0, 0.0, 5.0, -4.0], np.float32)
    x = np.array([3.0, -1.0, 5.0, -4.0], np.float32)
    s = oracle.verify(x, o)
    assert s["max_err"]["pwrrel"] == 0.0
    assert (s["max_err"]["abs"], s["max_err"]["idx"]) == (3.0, 0)
    x[3] = np.float32(-3.0)
    assert oracle.verify(x, o)["max_err"]["pwrrel"] == 0.25


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_fields_match_numpy_statistics(seed):
    rng = np.random.default_rng(seed)
    n = 50000 + 37 * seed
    o = (rng.standard_normal(n) * 3 + 100).astype(np.float32)
    o[rng.integers(0, n, 20)] = 0.0
    x = (o + rng.uniform(-1e-3, 1e-3, n).astype(np.float32)).astype(np.float32)
    s = oracle.verify(x, o)
    o64, x64 = o.astype(np.float64), x.astype(np.float64)
    assert s["odata"]["min"] == float(o.min()) and s["odata"]["max"] == float(o.max())
    assert s["xdata"]["min"] == float(x.min())
    assert s["odata"]["std"] == pytest.approx(np.std(o64), rel=1e-12)
    assert s["xdata"]["std"] == pytest.approx(np.std(x64), rel=1e-12)
    assert s["reduced"]["coeff"] == pytest.approx(np.corrcoef(o64, x64)[0, 1], rel=1e-12)
    err = np.abs(x - o)  # float32, as the reference's abserr
    assert s["reduced"]["MSE"] == pytest.approx(np.mean((err * err).astype(np.float64)), rel=1e-12)
    assert s["max_err"]["abs"] == float(err.max())
    assert s["max_err"]["idx"] == int(np.argmax(err))
    nz = o != 0
    assert s["max_err"]["pwrrel"] == float(np.max(err[nz] / np.abs(o[nz])))
    rng_o = float(o.max()) - float(o.min())
    assert s["reduced"]["PSNR"] == pytest.approx(20 * math.log10(rng_o) - 10 * math.log10(s["reduced"]["MSE"]),
                                                 rel=1e-14)


def test_identical_arrays_and_a_constant_field():
    o = np.linspace(0, 1, 64).astype(np.float32)
    s = oracle.verify(o, o)
    assert s["reduced"]["MSE"] == 0.0 and s["reduced"]["PSNR"] == math.inf
    assert s["max_err"]["abs"] == 0.0 and s["max_err"]["idx"] == 0
    c = np.full(16, 2.5, np.float32)
    s = oracle.verify(c + np.float32(0.5), c)
    assert s["odata"]["rng"] == 0.0 and s["odata"]["std"] == 0.0
    assert s["max_err"]["rel"] == math.inf and s["reduced"]["PSNR"] == -math.inf
    assert math.isnan(s["reduced"]["coeff"])
