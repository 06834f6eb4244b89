"""The C-ABI library loads and exports every symbol include/fz_capi.h declares; host-only
entry points (plans, sizes, argument validation, the scan's size query) work without a GPU."""
import ctypes
import os
import re

import pytest

import fzgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fz_capi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(fz_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_exports():
    assert _declared() == sorted(fzgpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = fzgpu.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert "gfx950" in fzgpu.version()


def test_plan_sizes_follow_the_reference_formulas():
    p = fzgpu.Plan(3600, 1800, 1, workspace=False)
    s = p.sizes
    assert (s.len, s.tiles, s.flag_words, s.byte_flags) == (6480000, 3164, 25313, 32 * 25313)
    assert s.tail_max_codes == 128
    assert s.payload_max_bytes == 4 * 25313 + 4096 * 3164
    assert s.workspace_bytes > 0


@pytest.mark.parametrize("args", [(0, 1, 1, 0, 0), (4, 4, 4, 10, 0), (8, 1, 1, 0, 7), (1, 1, 1, 2**33, 0)])
def test_plan_rejects_bad_geometry(args):
    with pytest.raises(fzgpu.FzError) as e:
        fzgpu.Plan(*args, workspace=False)
    assert e.value.status == 1


def test_tile_layout_knob_is_validated():
    p = fzgpu.Plan(100, 77, 1, workspace=False)
    p.set_tile_layout(1)
    p.set_tile_layout(0)
    for bad in (2, -1):
        with pytest.raises(fzgpu.FzError) as e:
            p.set_tile_layout(bad)
        assert e.value.status == 1


def test_inverse_form_knob_is_validated():
    p = fzgpu.Plan(64, 64, 64, workspace=False)
    for form in (fzgpu.INVERSE_WINDOWS, fzgpu.INVERSE_SPLIT, fzgpu.INVERSE_AUTO):
        p.set_inverse_form(form)
    for bad in (3, -1):
        with pytest.raises(fzgpu.FzError) as e:
            p.set_inverse_form(bad)
        assert e.value.status == 1


def test_scan_size_query_needs_no_device():
    n = ctypes.c_size_t(0)
    assert fzgpu.lib().fz_exclusive_sum_u32(None, ctypes.byref(n), None, None, 1 << 20, None) == 0
    assert n.value >= 256 + 8 * 256


def test_status_strings():
    L = fzgpu.lib()
    assert L.fz_status_string(4) == b"constant field (value range is 0)"
    assert L.fz_status_string(0) == b"success"


# ---------------------------------------------------------------- batched form (host side)

def test_batch_plan_sizes_follow_the_per_field_formulas():
    shapes = [(3600, 1800, 1), (5000, 1, 1), (33, 20, 17), (2047, 1, 1), (1000, 3, 1, 6000)]
    bp = fzgpu.BatchPlan(shapes, workspace=False)
    lens = [6480000, 5000, 33 * 20 * 17, 2047, 6000]
    assert bp.sizes.nfields == 5 and bp.sizes.total_len == sum(lens)
    assert bp.sizes.units == sum(n // 2048 + 1 for n in lens)
    assert bp.sizes.workspace_bytes >= 4096 * bp.sizes.units
    for i, s in enumerate(shapes):
        single = fzgpu.Plan(*s[:3], length=lens[i], workspace=False).sizes
        fs = bp.field_sizes(i)
        for k in ("len", "tiles", "flag_words", "payload_max_bytes", "tail_max_codes"):
            assert getattr(fs, k) == getattr(single, k), (i, k)


def test_batch_tuning_validation():
    bp = fzgpu.BatchPlan([(5000, 1, 1), (64, 64, 64)], workspace=False)
    t = bp.tuning
    assert t == {"range_chunk": 65536, "xcd_order": [1, 1, 1], "gather_waves": 16, "groups": 1,
                 "wave_tile": [0, 0, 0]}
    bp.set_tuning(range_chunk=16384, xcd_order=(0, 1, 0), gather_waves=4, groups=8, wave_tile=(1, 0, 1))
    assert bp.tuning == {"range_chunk": 16384, "xcd_order": [0, 1, 0], "gather_waves": 4, "groups": 8,
                         "wave_tile": [1, 0, 1]}
    parts16 = bp.sizes.range_parts
    for bad in [dict(range_chunk=1000), dict(gather_waves=3), dict(xcd_order=(2, 0, 0)), dict(groups=3),
                dict(groups=0), dict(wave_tile=(0, 2, 0))]:
        with pytest.raises(fzgpu.FzError) as e:
            bp.set_tuning(**bad)
        assert e.value.status == 1
    bp.set_tuning(range_chunk=262144)
    assert fzgpu.BatchPlan([(5000, 1, 1), (64, 64, 64)], workspace=False).sizes.range_parts >= bp.sizes.range_parts
    assert parts16 >= 2


@pytest.mark.parametrize("shapes", [[], [(0, 1, 1)], [(4, 4, 4, 10)], [(8, 1, 1)] * 4097])
def test_batch_plan_rejects_bad_input(shapes):
    with pytest.raises(fzgpu.FzError) as e:
        fzgpu.BatchPlan(shapes, workspace=False)
    assert e.value.status == 1


def test_batch_shapes_of_the_uneven_workload():
    from fzgpu import datagen as D
    spec = D.batch_fields()
    nd = [1 if f["dims"][1] == f["dims"][2] == 1 else (2 if f["dims"][2] == 1 else 3) for f in spec]
    assert (nd.count(1), nd.count(2), nd.count(3)) == (77, 102, 77)
    n = [f["dims"][0] * f["dims"][1] * f["dims"][2] for f in spec]
    assert min(n) >= 2 ** 16 * 0.9 and max(n) <= 2 ** 26 * 1.1
    assert {f["eb"] for f in spec} == {1e-2, 1e-3, 1e-4}
    bp = fzgpu.BatchPlan([f["dims"] for f in spec], workspace=False)
    assert bp.sizes.total_len == sum(n)


def test_verify_size_query_and_argument_checks_need_no_device():
    L = fzgpu.lib()
    n = ctypes.c_size_t(0)
    assert L.fz_verify_f32(None, ctypes.byref(n), None, None, None, 1 << 20, None) == 0
    big = n.value
    assert 0 < big <= 90 * 1024
    assert L.fz_verify_f32(None, ctypes.byref(n), None, None, None, 5, None) == 0
    assert 0 < n.value <= big
    assert L.fz_verify_f32(None, ctypes.byref(n), None, None, None, 0, None) == 1
    assert L.fz_verify_f32(None, None, None, None, None, 5, None) == 1
    # a non-null temp with missing arrays is rejected before any launch
    assert L.fz_verify_f32(ctypes.c_void_p(256), ctypes.byref(n), None, None, None, 5, None) == 1
    p = fzgpu.Plan(100, 77, 1, workspace=False)
    assert L.fz_decompress_verify_f32(p.handle, None, None, 1.0, None, 0, None, None, None, None) == 1


def test_outlier_mode_sizes_single_and_batched():
    p = fzgpu.Plan(100, 77, 1, mode=fzgpu.MODE_OUTLIER, workspace=False)
    assert p.sizes.side_max_bytes == 16 + 8 * 7700
    q = fzgpu.Plan(100, 77, 1, workspace=False)
    assert q.sizes.side_max_bytes == 0
    assert p.sizes.workspace_bytes > q.sizes.workspace_bytes  # unit counts and int32 deltas
    # an OUTLIER batch runs the single-field outlier launches in one workspace sized for its
    # largest field; each field's sizes are those of its single-field plan
    shapes = [(5000, 1, 1), (100, 77, 1), (33, 20, 17)]
    bp = fzgpu.BatchPlan(shapes, mode=fzgpu.MODE_OUTLIER, workspace=False)
    singles = [fzgpu.Plan(*s, mode=fzgpu.MODE_OUTLIER, workspace=False).sizes for s in shapes]
    assert bp.sizes.workspace_bytes == max(t.workspace_bytes for t in singles)
    for i, t in enumerate(singles):
        fs = bp.field_sizes(i)
        assert (fs.side_max_bytes, fs.payload_max_bytes, fs.tail_max_codes) == \
            (t.side_max_bytes, t.payload_max_bytes, t.tail_max_codes)
    half = fzgpu.BatchPlan(shapes, mode=fzgpu.MODE_OUTLIER, workspace=False, side_fraction=0.5)
    assert half.side_capacity(1) == 16 + 8 * 3850
    # the FIXED batch entry points refuse an OUTLIER plan, the OUTLIER ones a FIXED plan
    L = fzgpu.lib()
    fake = ctypes.c_void_p(1 << 20)
    arr = (ctypes.c_void_p * 3)(fake, fake, fake)
    ebs = (ctypes.c_double * 3)(1e-3, 1e-3, 1e-3)
    assert L.fz_batch_compress_f32(bp.handle, arr, ebs, fake, 1 << 40, arr, arr, fake, None) == 1
    fixed = fzgpu.BatchPlan(shapes, workspace=False)
    caps = (ctypes.c_uint64 * 3)(64, 64, 64)
    assert L.fz_batch_compress_outlier_f32(fixed.handle, arr, ebs, fake, 1 << 40, arr, arr, arr, caps, fake,
                                           None) == 1


# ---------------------------------------------------------------- archive input checks (host only)

def _oracle_payload(dims, mode=0, seed=5):
    import numpy as np

    import oracle as O
    from fzgpu import datagen as D
    x = D.waves(dims, seed).ravel().astype(np.float32)
    return O.compress(x, dims, 1e-3, mode)


@pytest.mark.parametrize("dims", [(1000, 21, 1), (6000, 1, 1), (20, 20, 20)])
def test_payload_check_accepts_oracle_payloads(dims):
    ref = _oracle_payload(dims)
    p = fzgpu.Plan(*dims, workspace=False)
    p.check_payload(ref["payload"], len(ref["tail"]))


def test_payload_check_refuses_inconsistent_payloads():
    import numpy as np
    dims = (1000, 21, 1)
    ref = _oracle_payload(dims)
    p = fzgpu.Plan(*dims, workspace=False)
    good, tl = ref["payload"].copy(), len(ref["tail"])
    W, T = p.sizes.flag_words, p.sizes.tiles
    assert tl == p.sizes.tail_max_codes > 0

    def refused(payload, tail_len=tl):
        with pytest.raises(fzgpu.FzError) as e:
            p.check_payload(payload, tail_len)
        assert e.value.status == 6

    refused(good[:-16])                    # one block short
    refused(good[: 4 * W - 4])             # shorter than the bitflags
    refused(np.concatenate([good, np.zeros(16, np.uint8)]))  # one block too many
    flipped = good.copy()
    b = int(np.flatnonzero(np.unpackbits(good[: 4 * W], bitorder="little") == 0)[0])
    flipped[b // 8] |= np.uint8(1 << (b % 8))  # a set flag with no block behind it
    refused(flipped)
    if W > 8 * T:  # a flag of the partial tile, which the reference never sets
        tail_flag = good.copy()
        tail_flag[4 * 8 * T] |= 1
        refused(np.concatenate([tail_flag, np.zeros(16, np.uint8)]))
    refused(good, tl - 1)                  # a short tail would make the decode read past it
    refused(good, 0)
    # LITERAL mode ignores the tail length
    lit = fzgpu.Plan(*dims, mode=fzgpu.MODE_LITERAL, workspace=False)
    ref_l = _oracle_payload(dims, mode=1)
    lit.check_payload(ref_l["payload"], 0)


@pytest.mark.parametrize("dims", [(300, 21, 1), (5000, 1, 1), (13, 12, 40)])
def test_payload_check_agrees_with_the_oracle_on_random_damage(dims):
    """Random damage to a valid archive (bitflag bits flipped, blocks cut or appended, bytes
    appended, the tail length changed): fz_payload_check accepts exactly the payloads the
    oracle's decompress accepts (fz_oracle.c decode_codes)."""
    import numpy as np

    import oracle as O
    ref = _oracle_payload(dims, seed=11)
    n = dims[0] * dims[1] * dims[2]
    p = fzgpu.Plan(*dims, workspace=False)
    good, tail = ref["payload"], ref["tail"]
    W = p.sizes.flag_words
    rng = np.random.default_rng(sum(dims))
    seen = {True: 0, False: 0}
    for trial in range(150):
        pay, tl = good.copy(), len(tail)
        kind = trial % 5
        if kind == 0:    # flip 1-3 bitflag bits
            for b in rng.integers(0, 32 * W, int(rng.integers(1, 4))):
                pay[b // 8] ^= np.uint8(1 << (b % 8))
        elif kind == 1:  # cut or append whole blocks
            k = int(rng.integers(1, 4))
            pay = pay[: max(0, pay.size - 16 * k)] if rng.integers(0, 2) else \
                np.concatenate([pay, rng.integers(0, 256, 16 * k).astype(np.uint8)])
        elif kind == 2:  # append a few stray bytes
            pay = np.concatenate([pay, np.zeros(int(rng.integers(1, 16)), np.uint8)])
        elif kind == 3:  # a flag moved: one set bit cleared and one clear bit set
            bits = np.unpackbits(pay[: 4 * W], bitorder="little")
            on, off = np.flatnonzero(bits), np.flatnonzero(bits == 0)
            if on.size and off.size:
                a, b = int(rng.choice(on)), int(rng.choice(off))
                pay[a // 8] ^= np.uint8(1 << (a % 8))
                pay[b // 8] ^= np.uint8(1 << (b % 8))
        else:            # the tail length
            tl = int(rng.integers(0, len(tail) + 3)) if len(tail) else 0
        t_arr = np.resize(tail, tl).astype(np.uint16) if tl else np.zeros(0, np.uint16)
        rc, _ = O.decompress(pay, t_arr if len(tail) else None, n, dims, ref["info"]["ebx2"])
        try:
            p.check_payload(pay, tl)
            ok = True
        except fzgpu.FzError as e:
            assert e.status == 6
            ok = False
        assert ok == (rc == 0), (trial, kind, rc)
        seen[ok] += 1
    assert seen[True] > 0 and seen[False] > 0


def test_side_check_follows_the_oracle_rules():
    import numpy as np
    dims = (100, 77, 1)
    p = fzgpu.Plan(*dims, mode=fzgpu.MODE_OUTLIER, workspace=False)
    xyz = 7700

    def side(idx_d, idx_r):
        h = np.array([len(idx_d), len(idx_r)], np.uint64).view(np.uint8)
        body = [np.asarray(idx_d, np.uint32).view(np.uint8), np.zeros(4 * len(idx_d), np.uint8),
                np.asarray(idx_r, np.uint32).view(np.uint8), np.zeros(4 * len(idx_r), np.uint8)]
        return np.concatenate([h] + body)

    p.check_side(side([], []))
    p.check_side(side([0, 5, xyz - 1], [7]))
    for bad in (side([xyz], []), side([], [xyz + 3]), side([1], [2])[:-4], np.zeros(8, np.uint8)):
        with pytest.raises(fzgpu.FzError) as e:
            p.check_side(bad)
        assert e.value.status == 6
    with pytest.raises(fzgpu.FzError) as e:  # only OUTLIER plans have a side channel
        fzgpu.Plan(*dims, workspace=False).check_side(side([], []))
    assert e.value.status == 1


def test_decompress_refuses_a_missing_tail_before_any_launch():
    L = fzgpu.lib()
    p = fzgpu.Plan(1000, 21, 1, workspace=False)  # 21000 floats: 10 tiles + 520 tail codes
    assert p.sizes.tail_max_codes == 520
    # aligned non-null stand-ins: the argument checks return before anything is dereferenced
    fake = ctypes.c_void_p(1 << 20)
    ws = ctypes.c_size_t(p.sizes.workspace_bytes)
    assert L.fz_decompress_f32(p.handle, fake, None, ctypes.c_float(0.5), fake, ws, fake, None) == 1


def _cli():
    path = os.path.join(ROOT, "fz-gpu_amd", "bin", "fz-gpu")
    if not os.path.exists(path):
        pytest.skip("fz-gpu CLI not built")
    return path


def test_cli_refuses_corrupt_archives_before_touching_the_gpu(tmp_path):
    """`fz-gpu -d` checks the header's sizes against the file and the payload's bitflags against
    its size on the host, so a damaged archive exits with "corrupt payload" (status 3) before any
    HIP call: this runs on a machine without a GPU."""
    import struct
    import subprocess

    import numpy as np
    dims = (1000, 21, 1)
    ref = _oracle_payload(dims)
    payload, tail = ref["payload"], ref["tail"]
    n = 21000

    def archive(pb=None, tl=None, body=None):
        pb = len(payload) if pb is None else pb
        tl = len(tail) if tl is None else tl
        h = struct.pack("<4sHBB3II QdfffI QQQ", b"FZMI", 1, 0, 2, *dims, 0, n, 1e-3,
                        float(ref["info"]["vmin"]), float(ref["info"]["vmax"]), float(ref["info"]["ebx2"]), 0,
                        pb, tl, 0)
        assert len(h) == 80
        body = payload.tobytes() + tail.astype(np.uint16).tobytes() if body is None else body
        f = tmp_path / "a.fz"
        f.write_bytes(h + body)
        return str(f)

    def run(a):
        return subprocess.run([_cli(), "-d", a, str(tmp_path / "out.f32")], capture_output=True, text=True,
                              timeout=60)

    cases = [archive(tl=len(tail) - 1, body=payload.tobytes() + tail[:-1].astype(np.uint16).tobytes()),
             archive(pb=len(payload) - 16, body=payload[:-16].tobytes() + tail.astype(np.uint16).tobytes())]
    for a in cases:
        r = run(a)
        assert r.returncode == 3 and "corrupt payload" in r.stderr, (r.returncode, r.stderr)
    # sizes the file cannot hold: refused by the reader
    r = run(archive(pb=1 << 40))
    assert r.returncode == 2 and "cannot read archive" in r.stderr, (r.returncode, r.stderr)
