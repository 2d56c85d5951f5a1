"""Pins the oracle's output-step functions (the post-analysis half of
write_ens_mpi, letkf/run/src/common_io.f90:863-1074):

  * ensemble_mesp (common.f90:505-596) against the reference's own compiled
    ensemble_mesp (oracle/_ref, the shipped FP32 build: nmem 128, nlev 75,
    nv3d 4, nv2d 3) bit for bit, and in FP64 against math.fsum means and
    spreads; nmem = 1 gives undef spreads, identical members zero spreads;
  * the single-precision global grids (gather_grd_mpi on one rank, the
    quasi-global wrap, the file's vertical order) against numpy's own
    reshape / flip / column copies of the same arrays.
"""
import math

import numpy as np
import pytest

import pyoracle as po


def _ens(rng, nij1, nlev, nmem, nv3d, nv2d, dtype):
    v3d = np.asfortranarray((rng.standard_normal((nij1, nlev, nmem, nv3d)) * 3.0 + 10.0).astype(dtype))
    v2d = np.asfortranarray((rng.standard_normal((nij1, nmem, nv2d)) * 0.4).astype(dtype))
    return v3d, v2d


def test_mesp_matches_reference_build():
    ref = po.ref()
    if ref is None:
        pytest.skip("reference library oracle/_ref not built (no reference sources here)")
    import ctypes as C
    nmem = ref.ref_nmem()
    assert ref.ref_real_kind() == 4
    nlev, nv3d, nv2d = 75, 4, 3   # common_setting.f90:39-41
    rng = np.random.default_rng(41)
    n = 37
    v3d, v2d = _ens(rng, n, nlev, nmem, nv3d, nv2d, np.float32)
    # a few points with identical members and with large offsets
    v3d[3] = v3d[3, :, :1, :]
    v3d[5] += 1.0e4
    want = [np.zeros((n, nlev, nv3d), np.float32, order="F") for _ in range(2)]
    want += [np.zeros((n, nv2d), np.float32, order="F") for _ in range(2)]
    F = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    ref.ref_ensemble_mesp(n, F(v3d), F(v2d), *(F(w) for w in want))
    got = po.ensemble_mesp(v3d, v2d, dtype=np.float32)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert np.all(got[1][3] < 1e-4)   # identical members: rounding of the mean only


def test_mesp_closed_forms():
    rng = np.random.default_rng(42)
    nij1, nlev, nmem, nv3d, nv2d = 23, 4, 9, 3, 2
    v3d, v2d = _ens(rng, nij1, nlev, nmem, nv3d, nv2d, np.float64)
    m3, s3, m2, s2 = po.ensemble_mesp(v3d, v2d)
    for i in range(nij1):
        for k in range(nlev):
            for v in range(nv3d):
                x = v3d[i, k, :, v]
                mean = math.fsum(x) / nmem
                assert abs(m3[i, k, v] - mean) <= 4e-16 * abs(mean) * nmem
                sd = math.sqrt(math.fsum((x - m3[i, k, v]) ** 2) / (nmem - 1))
                assert abs(s3[i, k, v] - sd) <= 1e-14 * sd
        for v in range(nv2d):
            x = v2d[i, :, v]
            assert abs(m2[i, v] - math.fsum(x) / nmem) <= 1e-15 * nmem
            assert abs(s2[i, v] - np.std(x, ddof=1)) <= 1e-14 * np.std(x, ddof=1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mesp_special_cases(dtype):
    rng = np.random.default_rng(43)
    v3d, v2d = _ens(rng, 11, 3, 1, 2, 1, dtype)
    m3, s3, m2, s2 = po.ensemble_mesp(v3d, v2d, undef=999.0, dtype=dtype)
    np.testing.assert_array_equal(m3, v3d[:, :, 0, :])
    assert np.all(s3 == 999.0) and np.all(s2 == 999.0)
    # identical members on a 1/64 lattice: exact mean, zero spread
    v3d, v2d = _ens(rng, 11, 3, 6, 2, 1, dtype)
    v3d = np.asfortranarray(np.repeat(np.round(v3d[:, :, :1, :] * 64) / 64, 6, axis=2).astype(dtype))
    m3, s3, _, _ = po.ensemble_mesp(v3d, v2d, dtype=dtype)
    np.testing.assert_array_equal(m3, v3d[:, :, 0, :])
    assert np.all(s3 == 0)


@pytest.mark.parametrize("wrap,file_order", [(False, False), (True, False), (True, True)])
def test_grid_pack_against_numpy(wrap, file_order):
    rng = np.random.default_rng(44)
    nlon, nlat, nlev, nf, nv3d, nv2d = 9, 5, 4, 3, 2, 2
    v3d, v2d = _ens(rng, nlon * nlat, nlev, nf, nv3d, nv2d, np.float64)
    g3, g2 = po.grid_pack(nlon, nlat, v3d, v2d, wrap=wrap, file_order=file_order)
    # numpy: (nij1, nlev, nf, nv3d) -> (nlon, nlat, nlev, nv3d, nf), Fortran order
    w3 = v3d.astype(np.float32).reshape(nlon, nlat, nlev, nf, nv3d, order="F").transpose(0, 1, 2, 4, 3)
    w2 = v2d.astype(np.float32).reshape(nlon, nlat, nf, nv2d, order="F").transpose(0, 1, 3, 2)
    w3, w2 = w3.copy(), w2.copy()
    if wrap:
        w3[-2:] = w3[:2]
        w2[-2:] = w2[:2]
    if file_order:
        w3 = w3[:, :, ::-1]
    np.testing.assert_array_equal(g3, w3)
    np.testing.assert_array_equal(g2, w2)
