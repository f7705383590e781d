"""The C-ABI library without a GPU: it loads, exports every function that
include/mara_gpu.h declares, its structs match the ctypes mirror, and
descriptor validation rejects what the reference rejects."""
import ctypes
import os
import re
import subprocess

import pytest

from mara2_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mara_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mara_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = runtime.load()
    names = declared_functions()
    assert len(names) >= 17
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", runtime.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mara_\w+)", out))
    assert set(names) <= exported


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text('''#include <stdio.h>
#include <stddef.h>
#include "mara_gpu.h"
#define F(T, m) printf(#T "." #m " %zu\\n", offsetof(T, m));
int main(void) {
  printf("mara_desc %zu\\n", sizeof(mara_desc));
  printf("mara_report %zu\\n", sizeof(mara_report));
  printf("mara_shape %zu\\n", sizeof(mara_shape));
  F(mara_shape, pencil) F(mara_shape, zplanes) F(mara_shape, rows)
  F(mara_desc, law) F(mara_desc, gamma) F(mara_desc, pressure_floor) F(mara_desc, recon)
  F(mara_desc, plm_theta) F(mara_desc, riemann) F(mara_desc, bc) F(mara_desc, cells)
  F(mara_desc, guards) F(mara_desc, lower) F(mara_desc, upper) F(mara_desc, rk_order)
  F(mara_desc, field_ct) F(mara_desc, precision) F(mara_desc, cooling_rate)
  F(mara_desc, geometry) F(mara_desc, log_radial)
  F(mara_report, status) F(mara_report, c2p_failures) F(mara_report, first_failure)
  F(mara_report, floor_applied) F(mara_report, eigen_failures) F(mara_report, superluminal)
  return 0; }
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in lines if l)
    assert int(got["mara_desc"]) == ctypes.sizeof(abi.MaraDesc)
    assert int(got["mara_report"]) == ctypes.sizeof(abi.MaraReport)
    assert int(got["mara_shape"]) == ctypes.sizeof(abi.MaraShape)
    for T, name in ((abi.MaraDesc, "mara_desc"), (abi.MaraReport, "mara_report"), (abi.MaraShape, "mara_shape")):
        for field, _ in T._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(T, field).offset, field


def test_error_strings_and_sizes():
    lib = runtime.load()
    assert lib.mara_error_string(abi.ERR_GUARD_TOO_SMALL).decode() == \
        "Boundary region of mesh data is smaller than the scheme's stencil"
    assert lib.mara_num_fields(abi.LAW_HYDRO) == 5 and lib.mara_num_fields(abi.LAW_RMHD) == 8
    for kw, nq in ((dict(law="mhd"), 8), (dict(law="mhd", cooling_rate=0.5), 9), (dict(law="mhd", cooling_rate=-1.0), 8),
                   (dict(law="hydro"), 5)):
        d = abi.make_desc(**kw)
        assert lib.mara_desc_num_fields(ctypes.byref(d)) == nq == abi.desc_num_fields(d), kw
    assert [lib.mara_stencil_size(r) for r in (0, 1, 2)] == [1, 2, 3]


@pytest.mark.parametrize("field,value,status", [
    ("rk_order", 4, abi.ERR_INVALID_ARGUMENT),      # MethodOfLinesTVD::setRungeKuttaOrder throws
    ("rk_order", 0, abi.ERR_INVALID_ARGUMENT),
    ("law", 7, abi.ERR_INVALID_ARGUMENT),
    ("recon", 5, abi.ERR_INVALID_ARGUMENT),
])
def test_create_rejects_bad_descriptors(field, value, status):
    lib = runtime.load()
    d = abi.make_desc()
    setattr(d, field, value)
    h = ctypes.c_void_p()
    assert lib.mara_block_create(ctypes.byref(d), ctypes.byref(h)) == status


def test_create_rejects_small_guards_and_hllc_mhd():
    lib = runtime.load()
    h = ctypes.c_void_p()
    d = abi.make_desc(recon="weno5", cells=(32, 32, 1), guards=(2, 2, 0))
    assert lib.mara_block_create(ctypes.byref(d), ctypes.byref(h)) == abi.ERR_GUARD_TOO_SMALL
    d = abi.make_desc(law="mhd", riemann="hllc")
    assert lib.mara_block_create(ctypes.byref(d), ctypes.byref(h)) == abi.ERR_UNSUPPORTED


def test_spherical_mesh_checks():
    """Geometry codes beyond the two meshes, flux-CT on a spherical mesh
    (CellCenteredFieldCT is the Cartesian construction), logarithmic radial
    edges from r = 0, and the entry points that do not take a spherical
    mesh: block groups and the no-allocation step."""
    lib = runtime.load()
    h = ctypes.c_void_p()
    d = abi.make_desc(geometry=5)
    assert lib.mara_block_create(ctypes.byref(d), ctypes.byref(h)) == abi.ERR_INVALID_ARGUMENT
    d = abi.make_desc(law="mhd", cells=(16, 16, 1), lower=(1, 0.2, 0), upper=(2, 2.9, 1), geometry="spherical")
    assert lib.mara_block_create(ctypes.byref(d), ctypes.byref(h)) == abi.ERR_UNSUPPORTED
    d = abi.make_desc(cells=(16, 1, 1), lower=(0, 0, 0), upper=(2, 3, 6), geometry="spherical", log_radial=True)
    assert lib.mara_block_create(ctypes.byref(d), ctypes.byref(h)) == abi.ERR_INVALID_ARGUMENT
    d = abi.make_desc(cells=(16, 16, 1), lower=(1, 0.2, 0), upper=(2, 2.9, 1), geometry="spherical")
    dims = (ctypes.c_int * 3)(2, 2, 1)
    assert lib.mara_group_create(ctypes.byref(d), dims, ctypes.byref(h)) == abi.ERR_UNSUPPORTED
    buf = (ctypes.c_double * 8)()
    assert lib.mara_advance_noalloc(ctypes.byref(d), buf, buf, ctypes.c_double(0.0), ctypes.c_double(0.1), None,
                                    None) == abi.ERR_UNSUPPORTED


def test_device_bytes_estimate():
    d = abi.make_desc(law="hydro", cells=(256, 256, 256), riemann="hllc")
    b = runtime.desc_device_bytes(d)
    assert 3 * 5 * 260 ** 3 * 8 <= b < 4 * 5 * 260 ** 3 * 8


def test_tuning_table_file_round_trip(tmp_path):
    """mara_tune_load / lookup / save without a GPU: entries are keyed by the
    descriptor's kernel law (cooling included), kernel selectors, cell counts
    and stage path (p2: 1-D pencils), a bad shape is refused."""
    d = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(96, 1, 1), field_ct=1)
    other = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(97, 1, 1), field_ct=1)
    cooled = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(96, 1, 1), field_ct=1,
                           cooling_rate=0.5)
    f = tmp_path / "tune.txt"
    f.write_text("# comment\n1 1 0 1 0 2 96 1 1 p2 : 512 8 2\n")
    runtime.tune_load(str(f))
    got = runtime.tune_lookup(d)
    assert got is not None and got.as_dict() == {"pencil": 512, "zplanes": 8, "rows": 2}
    assert runtime.tune_lookup(other) is None
    assert runtime.tune_lookup(cooled) is None
    out = tmp_path / "saved.txt"
    runtime.tune_save(str(out))
    assert "1 1 0 1 0 2 96 1 1 p2 : 512 8 2" in out.read_text().splitlines()
    bad = tmp_path / "bad.txt"
    bad.write_text("1 1 0 1 0 2 96 1 1 p2 : 300 8 2\n")
    with pytest.raises(runtime.MaraError):
        runtime.tune_load(str(bad))
    with pytest.raises(runtime.MaraError):
        runtime.tune_load(str(tmp_path / "missing.txt"))
