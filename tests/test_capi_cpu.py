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


def test_scan_size_query_needs_no_device():
    n = ctypes.c_size_t(0)
    assert fzgpu.lib().fz_exclusive_sum_u32(None, ctypes.byref(n), None, None, 1 << 20, None) == 0
    assert n.value >= 256 + 8 * 256


def test_status_strings():
    L = fzgpu.lib()
    assert L.fz_status_string(4) == b"constant field (value range is 0)"
    assert L.fz_status_string(0) == b"success"
