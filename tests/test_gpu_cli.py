"""The fz-gpu command line (fz.cu:226-237 drop-in) on the GPU: the three lines of
fz.cu:207-209 against the oracle's sizes, the archive round trip bit-exact against the
oracle's decompression, and the quality report of `fz-gpu -d <archive> <out> <original>`
against the oracle's verify_data (verify.hh:25-83, printed as quality_viewer.hh:53-95)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fz-gpu_amd", "bin", "fz-gpu")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(args):
    say("case cli", " ".join(os.path.basename(a) for a in args))
    p = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout


def _field(dims, seed):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    return (np.cumsum(rng.normal(size=n)) + 0.5 * np.sin(np.arange(n) / 50.0)).astype(np.float32)


@pytest.mark.parametrize("dims,eb,literal", [((600, 300, 1), 1e-3, False), ((48, 40, 36), 1e-4, False),
                                              ((70001, 1, 1), 1e-2, True)])
def test_cli_lines_archive_and_quality_report(tmp_path, dims, eb, literal):
    x = _field(dims, 3)
    src = tmp_path / "field.f32"
    x.tofile(src)
    arc = tmp_path / "field.fzmi"
    mode = O.MODE_LITERAL if literal else O.MODE_FIXED
    args = [str(src), *map(str, dims), repr(eb), "-o", str(arc)] + (["--literal"] if literal else [])
    lines = _run(args).splitlines()
    ref = O.compress(x, dims, eb, mode)
    assert ref["status"] == 0
    csize = ref["info"]["payload_bytes"]
    assert lines[0] == f"original size: {4 * x.size}"
    assert lines[1] == f"compressed size: {csize}"
    assert lines[2] == "compression ratio: %f" % float(np.float32(4 * x.size) / np.float32(csize))

    out = tmp_path / "out.f32"
    assert _run(["-d", str(arc), str(out)]) == ""
    rc, ref_out = O.decompress(ref["payload"], ref["tail"] if mode == O.MODE_FIXED else None, x.size, dims,
                               ref["info"]["ebx2"])
    assert rc == 0
    got = np.fromfile(out, dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), ref_out.view(np.uint32))

    rep = _run(["-d", str(arc), str(out), str(src)])
    assert "quality metrics (using GPU checker):" in rep
    s = O.verify(ref_out, x)
    nums = {}
    for line in rep.splitlines():
        m = re.match(r"\s+(origin|eb-lossy|max-error|metrics|auto)\s+(.*)$", line)
        if m:
            nums[m.group(1)] = [float(v) for v in m.group(2).split()]
    assert nums["origin"] == pytest.approx([s["odata"]["min"], s["odata"]["max"], s["odata"]["rng"],
                                            s["odata"]["std"]], rel=1e-7)
    assert nums["max-error"] == pytest.approx([s["max_err"]["abs"], s["max_err"]["idx"], s["max_err"]["pwrrel"],
                                               s["max_err"]["rel"]], rel=1e-7)
    assert nums["metrics"][1:] == pytest.approx([s["reduced"]["NRMSE"], s["reduced"]["coeff"],
                                                 s["reduced"]["PSNR"]], rel=1e-7)
    assert nums["metrics"][0] == pytest.approx(4 * x.size / os.path.getsize(arc), rel=1e-7)
    lag1 = O.verify(x[:-1], x[1:])["reduced"]["coeff"]
    lag2 = O.verify(x[:-2], x[2:])["reduced"]["coeff"]
    assert nums["auto"] == pytest.approx([lag1, lag2], abs=1e-6)


def _run_rc(args):
    say("case cli", " ".join(os.path.basename(a) for a in args))
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("dims", [(600, 300, 1), (30, 40, 50)])
def test_cli_outlier_mode_archive_round_trip(tmp_path, dims):
    x = _field(dims, 8)
    rng = np.random.default_rng(2)
    x[rng.integers(0, x.size, 50)] = np.nan
    x[rng.integers(0, x.size, 50)] = np.float32(5e6)
    src = tmp_path / "f.f32"
    x.tofile(src)
    arc = tmp_path / "f.fzmi"
    lines = _run([str(src), *map(str, dims), "1e-6", "-o", str(arc), "--outliers"]).splitlines()
    ref = O.compress_outlier(x, dims, 1e-6)
    assert ref["status"] == 0 and ref["info"]["overflow"] > 0
    assert lines[1] == f"compressed size: {ref['info']['payload_bytes']}"
    assert os.path.getsize(arc) == 80 + ref["info"]["payload_bytes"] + 2 * ref["info"]["tail_len"] + \
        ref["info"]["side_bytes"]
    out = tmp_path / "o.f32"
    assert _run(["-d", str(arc), str(out)]) == ""
    rc, ref_out = O.decompress_outlier(ref["payload"], ref["tail"], ref["side"], x.size, dims, ref["info"]["ebx2"],
                                       ref["info"]["vmin"])
    assert rc == 0
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), ref_out.view(np.uint32))


def test_cli_constant_field_needs_the_outlier_mode(tmp_path):
    c = np.full(20000, 7.5, np.float32)
    src = tmp_path / "c.f32"
    c.tofile(src)
    p = _run_rc([str(src), "200", "100", "1", "1e-3"])
    assert p.returncode != 0 and "constant" in p.stderr
    arc = tmp_path / "c.fzmi"
    _run([str(src), "200", "100", "1", "1e-3", "-o", str(arc), "--outliers"])
    out = tmp_path / "c.out"
    _run(["-d", str(arc), str(out)])
    assert np.array_equal(np.fromfile(out, dtype=np.float32), c)
