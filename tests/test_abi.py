"""The C ABI library loads on a host without a GPU and exports every entry
point include/fluidphysics.h declares; the ctypes host binds the same set.
No compute calls (no device here)."""
import ctypes
import json
import os
import re
import subprocess

import pytest

from fluidphysics_amd import _abi

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def _declared():
    src = open(os.path.join(ROOT, "include", "fluidphysics.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fp_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_reference_surface():
    names = _declared()
    for fn in ["fp_create", "fp_destroy", "fp_add_solid_cube", "fp_step", "fp_cut", "fp_set_grabs",
               "fp_emit", "fp_pick_solid", "fp_reset_solid", "fp_clear_fluid", "fp_fluid_count",
               "fp_read_state", "fp_read_grid", "fp_read_constraints"]:
        assert fn in names


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fp_[a-z_0-9]+)", out))
    assert set(_declared()) <= exported


def test_ctypes_host_binds_exactly_the_header():
    assert sorted(_abi.SIGNATURES) == _declared()


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_params_default_and_error_paths_without_device():
    lib = _abi.load_library()
    p = _abi.FpParams()
    lib.fp_params_default(ctypes.byref(p))
    ref = dict(_abi.REFERENCE_DEFAULTS)
    for k, v in ref.items():
        assert getattr(p, k) == v, k
    # argument errors are reported without touching a device
    assert lib.fp_create(ctypes.byref(p), -1, 0, ctypes.byref(ctypes.c_void_p())) == _abi.FP_EINVAL
    assert b"max_fluid" in lib.fp_last_error(None)
    bad = _abi.make_params(cell_size=0.1, fluid_h=0.19)
    assert lib.fp_create(ctypes.byref(bad), 16, 0, ctypes.byref(ctypes.c_void_p())) == _abi.FP_EINVAL
    assert lib.fp_step(None, 0.01, 1) == _abi.FP_EINVAL


def test_params_struct_layout_matches_header():
    # 24 doubles, 4 int32, 4 doubles, 2 int32, 1 double
    assert ctypes.sizeof(_abi.FpParams) == 24 * 8 + 4 * 4 + 4 * 8 + 2 * 4 + 8


def _c_layout(struct, fields, tmp_path):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fluidphysics.h"', 'int main(void) {',
             '  printf("size %%zu\\n", sizeof(%s));' % struct]
    lines += ['  printf("%s %%zu\\n", offsetof(%s, %s));' % (f, struct, f) for f in fields]
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    return {k: int(v) for k, v in (ln.split() for ln in out if ln)}


@pytest.mark.parametrize("struct,cls", [("fp_params", "FpParams"), ("fp_info", "FpInfo"),
                                        ("fp_pass_stat", "FpPassStat"), ("fp_batch_info", "FpBatchInfo")])
def test_struct_offsets_match_the_c_compiler(struct, cls, tmp_path):
    """Every field of the ctypes mirror sits where gcc puts it."""
    c = getattr(_abi, cls)
    fields = [f for f, _ in c._fields_]
    lay = _c_layout(struct, fields, tmp_path)
    assert lay.pop("size") == ctypes.sizeof(c)
    for f in fields:
        assert lay[f] == getattr(c, f).offset, f


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_scene_bytes_bounds_the_cube_by_the_csr_width():
    """2 x constraints of an n^3 lattice must fit an int32 CSR: n = 436 is the
    largest cube (2C = 2,144,678,760 < INT_MAX); n = 437 gives 2,159,493,176
    and is refused before anything is sized or allocated."""
    lib = _abi.load_library()
    p = _abi.FpParams()
    lib.fp_params_default(ctypes.byref(p))
    out = ctypes.c_int64(0)
    c = lambda n: 3 * n * n * (n - 1) + 6 * n * (n - 1) ** 2 + 4 * (n - 1) ** 3
    assert 2 * c(436) < 2 ** 31 - 1 < 2 * c(437)
    assert lib.fp_scene_bytes(ctypes.byref(p), 0, 436, ctypes.byref(out)) == _abi.FP_OK
    assert out.value > 2 * c(436) * 4  # at least the per-entry rest lengths
    assert lib.fp_scene_bytes(ctypes.byref(p), 0, 437, ctypes.byref(out)) == _abi.FP_EINVAL
    # the capacity bound (n^3 + maxFluid < 2^27) still holds
    assert lib.fp_scene_bytes(ctypes.byref(p), (1 << 27) - 8, 2, ctypes.byref(out)) == _abi.FP_EINVAL


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_noalloc_grid_refuses_sizes_beyond_its_array_without_a_device():
    """fp_build_grid_noalloc checks its sizes before it touches a device: n
    above FP_GRID_NOALLOC_MAX_N, a table above FP_GRID_NOALLOC_MAX_T or not a
    power of two, a bad cell size or a NULL pointer give FP_EINVAL."""
    lib = _abi.load_library()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every call below is refused first
    call = lambda n, t, cs=0.22, p=fake: lib.fp_build_grid_noalloc(p, fake, n, t, cs, fake, fake, None)
    assert call((1 << 20) + 1, 1 << 21) == _abi.FP_EINVAL
    assert b"n out of range" in lib.fp_last_error(None)
    assert call(-1, 1024) == _abi.FP_EINVAL
    assert call(100, (1 << 21) * 2) == _abi.FP_EINVAL
    assert call(100, 1000) == _abi.FP_EINVAL
    assert b"power of two" in lib.fp_last_error(None)
    assert call(100, 1024, cs=0.0) == _abi.FP_EINVAL
    assert call(100, 1024, cs=float("nan")) == _abi.FP_EINVAL
    assert call(100, 1024, p=None) == _abi.FP_EINVAL


def test_tuning_cache_file_round_trip_and_rejects(tmp_path):
    """fp_tune_cache_save / _load / _reset need no device: an empty cache
    round-trips, a written file loads back, and malformed files or
    out-of-range values are refused without merging anything."""
    import fluidphysics_amd as F
    F.tune_cache_reset()
    p = tmp_path / "tune.json"
    F.tune_cache_save(str(p))
    d = json.loads(p.read_text())
    assert d == {"version": 1, "contact_lanes": {}, "launch_shape": {}}
    good = {"version": 1, "contact_lanes": {"1099528577282": 8}, "launch_shape": {"16843266": 3}}
    p.write_text(json.dumps(good))
    F.tune_cache_load(str(p))
    q = tmp_path / "back.json"
    F.tune_cache_save(str(q))
    assert json.loads(q.read_text()) == good
    for bad in ['{"version": 2, "contact_lanes": {}, "launch_shape": {}}',
                '{"version": 1, "contact_lanes": {"12": 7}}',
                '{"version": 1, "launch_shape": {"12": 6}}',
                '{"version": 1, "contact_lanes": {"x1": 8}}',
                '{"version": 1, "other": {}}',
                '{"version": 1} trailing',
                'not json']:
        p.write_text(bad)
        with pytest.raises(F.FpError):
            F.tune_cache_load(str(p))
    with pytest.raises(F.FpError):
        F.tune_cache_load(str(tmp_path / "missing.json"))
    F.tune_cache_save(str(q))
    assert json.loads(q.read_text()) == good   # nothing merged by the refused files
    F.tune_cache_reset()
    F.tune_cache_save(str(q))
    assert json.loads(q.read_text())["contact_lanes"] == {}
