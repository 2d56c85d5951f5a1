"""Host-side checks of the C ABI (no GPU compute): the library loads and
exports every function include/letkf_engine.h declares, the ctypes and
Fortran mirrors of the structs match the C layout, defaults equal the
reference PARAMETERs, and argument errors come back as status codes."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from sbpom_letkf_v2_amd import _lib, engine, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "letkf_engine.h")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(letkf_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 10
    L = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "lay.c"
    src.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "letkf_engine.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(letkf_params_t), sizeof(letkf_grid_t), sizeof(letkf_obs_t), sizeof(letkf_stats_t));
  printf("%zu %zu %zu %zu\n", offsetof(letkf_params_t, pi), offsetof(letkf_params_t, smax), offsetof(letkf_params_t, stage_timing), offsetof(letkf_params_t, workspace_bytes));
  printf("%zu %zu\n", offsetof(letkf_obs_t, nobsgrd), offsetof(letkf_stats_t, ms_total));
  printf("%zu %zu %zu %zu\n", sizeof(letkf_xtoy_grid_t), offsetof(letkf_xtoy_grid_t, hnosea),
         sizeof(letkf_setobs_cfg_t), offsetof(letkf_setobs_cfg_t, qc_max));
  printf("%zu %zu\n", sizeof(letkf_obs_set_t), offsetof(letkf_obs_set_t, nobs_out));
  return 0;
}''')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = [int(x) for x in out]
    want = [C.sizeof(_lib.Params), C.sizeof(_lib.Grid), C.sizeof(_lib.Obs), C.sizeof(_lib.Stats),
            _lib.Params.pi.offset, _lib.Params.smax.offset, _lib.Params.stage_timing.offset,
            _lib.Params.workspace_bytes.offset,
            _lib.Obs.nobsgrd.offset, _lib.Stats.ms_total.offset,
            C.sizeof(_lib.XtoyGrid), _lib.XtoyGrid.hnosea.offset, C.sizeof(_lib.SetObsCfg),
            _lib.SetObsCfg.qc_max.offset, C.sizeof(_lib.ObsSet), _lib.ObsSet.nobs_out.offset]
    assert got == want


def test_defaults_are_reference_parameters():
    L = _lib.lib()
    for rk in (4, 8):
        p = _lib.Params()
        assert L.letkf_params_default(C.byref(p), rk) == 0
        R = (lambda x: float(np.float32(x))) if rk == 4 else (lambda x: x)
        assert p.real_kind == rk and p.iv3d_t == 3 and p.iv3d_s == 4 and p.id_z_obs == 2567
        assert p.sigma_obs == R(300e3) and p.sigma_obsv == R(100.0)
        assert p.alpha_rtpp == R(0.9) and p.alpha_rtps == 0.0 and p.cov_infl_mul == 1.0
        assert p.tmin == R(-1.8) and p.smax == R(50.0) and p.re == R(6371.3e3)
        c = R(2.0 * np.sqrt(10.0 / 3.0))
        dz = float(np.float32(c) * np.float32(R(300e3))) if rk == 4 else c * 300e3
        assert p.dist_zero == dz
    assert L.letkf_params_default(C.byref(_lib.Params()), 3) == -1


def test_argument_errors_are_status_codes():
    L = _lib.lib()
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=4, nobs=10, seed=1)
    p = engine.make_params(pr)
    h = C.c_void_p()
    bad = _lib.Params.from_buffer_copy(p)
    bad.nmem = 1
    assert L.letkf_create(C.byref(bad), C.byref(h)) == -1
    assert b"nmem" in L.letkf_last_error()
    bad = _lib.Params.from_buffer_copy(p)
    bad.real_kind = 2
    assert L.letkf_create(C.byref(bad), C.byref(h)) == -1
    bad = _lib.Params.from_buffer_copy(p)
    bad.nv3d, bad.nv2d = 7, 2
    assert L.letkf_create(C.byref(bad), C.byref(h)) == -1
    bad = _lib.Params.from_buffer_copy(p)
    bad.precision = 9
    assert L.letkf_create(C.byref(bad), C.byref(h)) == -1
    # every named precision has a value the C header accepts
    assert set(engine.PRECISIONS.values()) == {0, 1, 2, 3, 4, 5, 6}
    hdr = open(HEADER).read()
    for name, v in (("LETKF_PREC_FP32_PANEL", 3), ("LETKF_PREC_FP32X3", 4), ("LETKF_PREC_FP32_DUAL64", 5),
                    ("LETKF_PREC_BF16X3_TRI32", 6)):
        assert re.search(rf"{name} = {v}\b", hdr), name
    # and every named form
    assert set(engine.FORMS.values()) == {0, 1, 2}
    for name, v in (("LETKF_FORM_PLAIN", 0), ("LETKF_FORM_BATCHED", 1), ("LETKF_FORM_BATCHED_QL", 2)):
        assert re.search(rf"{name} = {v}\b", hdr), name
    bad = _lib.Params.from_buffer_copy(p)
    bad.form = 3
    assert L.letkf_create(C.byref(bad), C.byref(h)) == -1
    # NULL handle / arrays
    g, o = _lib.Grid(), _lib.Obs()
    assert L.letkf_da_main(None, C.byref(g), C.byref(o), None, None, None, None) == -1
    assert L.letkf_destroy(None) == 0
    # every device entry point rejects a NULL handle before any HIP call
    # (this also checks that the ctypes argument lists match the C arity)
    cfg, out = _lib.SetObsCfg(), _lib.ObsSet()
    L.letkf_setobs_default(C.byref(cfg))
    assert L.letkf_set_obs_device(None, C.byref(cfg), 8, 6, 0, None, *([None] * 11), C.byref(out)) == -1
    assert L.letkf_bin_obs(None, 8, 6, None, None, 0, None, None, None, None, None, None, None) == -1
    assert L.letkf_aoei(None, 0, None, None, None) == -1
    assert L.letkf_get_nobsl(None, None, None) == -1
    assert L.letkf_core_batched_noalloc(None, 0, *([None] * 9)) == -1
    assert L.letkf_core_batched(None, 0, None, None, None, None, None, None, None) == -1
    assert L.letkf_da_main_device(None, C.byref(g), C.byref(o), None, None, None, None, None) == -1
    assert L.letkf_da_adjoint_device(None, C.byref(g), C.byref(o), *([None] * 8)) == -1
    xg = _lib.XtoyGrid(nlon=8, nlat=6, nlev=3, nv3d=4, nv2d=1, real_kind=8)
    assert L.letkf_trans_xtoy_device(C.byref(xg), 1, *([None] * 6), 2, None, None, None, None, None,
                                     None) == -1
    assert L.letkf_lpfgm_core_batched(None, 3, 1.0, 100, 0, 1, None, 0, *([None] * 7)) == -1
    lc = _lib.LpfgmCfg()
    assert L.letkf_da_main_lpfgm_device(None, C.byref(g), C.byref(o), C.byref(lc), *([None] * 5)) == -1
    assert L.letkf_da_main_lpfgm(None, C.byref(g), C.byref(o), C.byref(lc), *([None] * 4)) == -1
    # the output-step entries check kinds and sizes before any HIP call
    assert L.letkf_ensemble_mesp_device(3, 10, 2, 4, 1, 1, 999.0, *([None] * 7)) == -1
    assert L.letkf_ensemble_mesp_device(8, 10, 2, 0, 1, 1, 999.0, *([None] * 7)) == -1
    assert L.letkf_ensemble_mesp_device(8, 0, 2, 4, 1, 1, 999.0, *([None] * 7)) == 0
    assert L.letkf_grid_pack_device(8, 2, 10, 3, 1, 1, 1, 1, 0, *([None] * 7)) == -1   # wrap, nlon < 4
    assert L.letkf_grid_pack_device(8, 8, 10, 3, 1, 1, 1, 0, 0, *([None] * 7)) == -1   # NULL arrays
    assert L.letkf_grid_pack_device(8, 8, 10, 3, 1, 1, 0, 1, 0, *([None] * 7)) == 0    # no fields
    with pytest.raises(_lib.EngineError):
        _lib.check(-1)


def test_create_without_device_fails_cleanly():
    L = _lib.lib()
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=4, nobs=10, seed=1)
    p = engine.make_params(pr)
    h = C.c_void_p()
    rc = L.letkf_create(C.byref(p), C.byref(h))
    if rc == 0:                      # a device is present: release it again
        assert L.letkf_destroy(h) == 0
    else:
        assert rc == -2 and L.letkf_last_error()


@pytest.mark.skipif(shutil.which("amdflang") is None or
                    not os.path.isdir("/root/reference/letkf/run/src"),
                    reason="needs amdflang and the reference's common_setting.f90")
def test_fortran_dropin_compiles_against_reference_module(tmp_path):
    """letkf_da_main_gpu compiles against the reference's own common_setting
    module (the module state da_main reads) and a driver calling it links."""
    ref = "/root/reference/letkf/run/src/common_setting.f90"
    mod = os.path.join(ROOT, "sbpom_letkf_v2_amd", "fortran", "letkf_engine_mod.f90")
    drv = tmp_path / "drv.f90"
    drv.write_text("""
program drv
  use common_setting
  use letkf_da_gpu
  implicit none
  real(r_size), allocatable :: f3(:,:,:,:), f2(:,:,:), a3(:,:,:,:), a2(:,:,:)
  real(8) :: draws(nmem, nmonte)
  nij1 = 1
  allocate(f3(nij1,nlev,nmem,nv3d), f2(nij1,nmem,nv2d), a3(nij1,nlev,nmem,nv3d), a2(nij1,nmem,nv2d))
  if (.false.) call letkf_da_main_gpu(f3, f2, a3, a2)
  if (.false.) call letkf_da_main_gpu(f3, f2, a3, a2, rand=draws)
end program drv
""")
    subprocess.run(["amdflang", "-c", ref, "-o", str(tmp_path / "cs.o")], check=True, cwd=tmp_path)
    subprocess.run(["amdflang", "-c", mod, "-o", str(tmp_path / "m.o")], check=True, cwd=tmp_path)
    subprocess.run(["amdflang", "-c", str(drv), "-o", str(tmp_path / "d.o")], check=True, cwd=tmp_path)
    subprocess.run(["amdflang", str(tmp_path / "d.o"), str(tmp_path / "m.o"), str(tmp_path / "cs.o"),
                    "-L", os.path.dirname(_lib.LIB_PATH), "-lletkf_engine",
                    "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH), "-o", str(tmp_path / "drv")],
                   check=True, cwd=tmp_path)
    assert (tmp_path / "drv").exists()
