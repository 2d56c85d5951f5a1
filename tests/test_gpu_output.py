"""GPU parity of the output step after da_main (write_ens_mpi,
letkf/run/src/common_io.f90:863-1074) against the oracle on the same inputs:
letkf_ensemble_mesp_device (ensemble_mesp, common.f90:505-596) and
letkf_grid_pack_device (gather_grd_mpi on one rank, the quasi-global wrap,
the file's vertical order).  The bar is bit-exact: both sides run the
member loops in member order with the same operations, and the pack is a
cast and a permutation."""
import numpy as np
import pytest
import torch

import pyoracle as po
from sbpom_letkf_v2_amd import engine, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ens(rng, nij, nlev, nmem, nv3d, nv2d, dtype):
    v3d = np.asfortranarray((rng.standard_normal((nij, nlev, nmem, nv3d)) * 3.0 + 10.0).astype(dtype))
    v2d = np.asfortranarray((rng.standard_normal((nij, nmem, nv2d)) * 0.4).astype(dtype))
    return v3d, v2d


def _dev3(a):   # Fortran (nij, nlev, nmem, nv3d) -> C (nv3d, nmem, nlev, nij) on the GPU
    return torch.from_numpy(np.ascontiguousarray(a.transpose(3, 2, 1, 0))).cuda()


def _dev2(a):   # Fortran (nij, nmem, nv2d) -> C (nv2d, nmem, nij)
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nmem", [1, 13, 100])
def test_ensemble_mesp_bit_exact(dtype, nmem):
    _say(f"case mesp {np.dtype(dtype).name} nmem={nmem}")
    rng = np.random.default_rng(51 + nmem)
    nij, nlev, nv3d, nv2d = 1000, 7, 5, 1
    v3d, v2d = _ens(rng, nij, nlev, nmem, nv3d, nv2d, dtype)
    v3d[::97] += 1.0e4   # large offsets: the deviations carry the mean's rounding
    want = po.ensemble_mesp(v3d, v2d, undef=999.0, dtype=dtype)
    got = engine.ensemble_mesp_device(_dev3(v3d), _dev2(v2d), undef=999.0)
    torch.cuda.synchronize()
    m3, s3, m2, s2 = (g.cpu().numpy() for g in got)
    np.testing.assert_array_equal(m3, want[0].transpose(2, 1, 0))
    np.testing.assert_array_equal(s3, want[1].transpose(2, 1, 0))
    np.testing.assert_array_equal(m2, want[2].T)
    np.testing.assert_array_equal(s2, want[3].T)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("wrap,file_order", [(False, False), (True, False), (True, True)])
def test_grid_pack_bit_exact(dtype, wrap, file_order):
    _say(f"case grid_pack {np.dtype(dtype).name} wrap={wrap} file_order={file_order}")
    rng = np.random.default_rng(61)
    nlon, nlat, nlev, nf, nv3d, nv2d = 37, 23, 6, 5, 4, 2
    v3d, v2d = _ens(rng, nlon * nlat, nlev, nf, nv3d, nv2d, dtype)
    w3, w2 = po.grid_pack(nlon, nlat, v3d, v2d, wrap=wrap, file_order=file_order, dtype=dtype)
    g3, g2 = engine.grid_pack_device(nlon, nlat, _dev3(v3d), _dev2(v2d), wrap=wrap, file_order=file_order)
    torch.cuda.synchronize()
    # (nf, nv3d, nlev, nlat, nlon) C order == (nlon, nlat, nlev, nv3d, nf) Fortran order
    np.testing.assert_array_equal(g3.cpu().numpy(), w3.transpose(4, 3, 2, 1, 0))
    np.testing.assert_array_equal(g2.cpu().numpy(), w2.transpose(3, 2, 1, 0))
    # one field (a mean or a spread): (nv3d, nlev, nij) / (nv2d, nij)
    m3 = _dev3(v3d)[:, 2].contiguous()
    m2 = _dev2(v2d)[:, 2].contiguous()
    h3, h2 = engine.grid_pack_device(nlon, nlat, m3, m2, wrap=wrap, file_order=file_order)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h3[0].cpu().numpy(), w3[..., 2].transpose(3, 2, 1, 0))
    np.testing.assert_array_equal(h2[0].cpu().numpy(), w2[..., 2].transpose(2, 1, 0))


def test_output_errors():
    v = torch.zeros((1, 2, 3, 20), dtype=torch.float64, device="cuda")
    w = torch.zeros((1, 2, 20), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):   # the wrap needs nlon >= 4
        engine.grid_pack_device(2, 10, v, w, wrap=True)
    engine.grid_pack_device(2, 10, v, w, wrap=False)
    with pytest.raises(ValueError):
        engine.grid_pack_device(4, 4, v, w)


def test_output_chain_hxa():
    """The obs-space part of write_ens_mpi in the reference's shipped REAL(4)
    build: the analysis members and the analysis mean/spread grids are
    packed (wrap, model order) and go through Trans_XtoY (common_io.f90:
    966-989) -- letkf_ensemble_mesp_device -> letkf_grid_pack_device ->
    letkf_trans_xtoy_device against po.ensemble_mesp -> po.grid_pack ->
    po.trans_xtoy, bit for bit."""
    _say("case output chain H(x^a)")
    dt = np.float32
    nlon, nlat, nlev, nmem = 40, 30, 8, 12
    pr = synth.make_problem(nlon=nlon, nlat=nlat, nlev=nlev, nmem=nmem, nobs=10, seed=7, real_kind=4,
                            land_frac=0.05)
    pn = pr.numpy()
    rng = np.random.default_rng(71)
    lon2 = np.ascontiguousarray(pn.lon1.reshape(nlat, nlon))
    lat2 = np.ascontiguousarray(pn.lat1.reshape(nlat, nlon))
    depth3 = np.ascontiguousarray(pn.depth1.T.reshape(nlev, nlat, nlon))
    fsm = np.ascontiguousarray((pn.phi1.reshape(nlat, nlon) > 1.0).astype(dt))
    n = 2000
    olon = rng.uniform(lon2[0, 0], lon2[0, -1], n).astype(dt)
    olat = rng.uniform(lat2[0, 0], lat2[-1, 0], n).astype(dt)
    elm = rng.choice([synth.ID_U, synth.ID_V, synth.ID_T, synth.ID_S, synth.ID_Z], n).astype(np.int32)
    olev = np.where((elm == synth.ID_Z) | (rng.uniform(size=n) < 0.3), 0.0,
                    rng.uniform(depth3[1].max(), 0.0, n)).astype(dt)
    idx, idy, _, _ = po.bin_obs(lon2.T, lat2.T, olon, olat, dtype=dt)
    a3, a2 = pn.fcst3d, pn.fcst2d          # (nij1, nlev, nmem, nv3d), (nij1, nmem, nv2d): an ensemble
    # oracle side
    m3, s3, m2, s2 = po.ensemble_mesp(a3, a2, dtype=dt)
    w3, w2 = po.grid_pack(nlon, nlat, a3, a2, wrap=True, dtype=dt)
    f3 = np.asfortranarray(np.stack([m3, s3], axis=2))      # (nij1, nlev, 2, nv3d)
    f2 = np.asfortranarray(np.stack([m2, s2], axis=1))
    x3, x2 = po.grid_pack(nlon, nlat, f3, f2, wrap=True, dtype=dt)
    g = po.OracleXtoY(nlon, nlat, nlev, 4, 1, synth.ID_U, synth.ID_V, synth.ID_T, synth.ID_S, synth.ID_Z,
                      1, 2, 3, 4, 1, 0.0, 999.0)
    H = lambda v3, v2: po.trans_xtoy(g, elm, idx, idy, olon, olat, olev, lon2.T, lat2.T, depth3.T, fsm.T,
                                     v3, v2, dtype=dt)   # v3 (nlon,nlat,nlev,nv3d), v2 (nlon,nlat,nv2d)
    want_hxa = np.stack([H(w3[..., m], w2[..., m]) for m in range(nmem)])
    want_mean = H(x3[..., 0], x2[..., 0])
    want_sprd = H(x3[..., 1], x2[..., 1])
    # device side
    prd = pr.to("cuda")
    gm3, gs3, gm2, gs2 = engine.ensemble_mesp_device(prd.fcst3d, prd.fcst2d)
    p3, p2 = engine.grid_pack_device(nlon, nlat, prd.fcst3d, prd.fcst2d, wrap=True)
    q3, q2 = engine.grid_pack_device(nlon, nlat, torch.stack([gm3, gs3], 1).contiguous(),
                                     torch.stack([gm2, gs2], 1).contiguous(), wrap=True)
    cu = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    args = (cu(lon2), cu(lat2), cu(depth3), cu(fsm), cu(elm), cu(idx), cu(idy), cu(olon), cu(olat), cu(olev))
    hxa, _ = engine.trans_xtoy_device(*args, p3, p2)
    hms, _ = engine.trans_xtoy_device(*args, q3, q2)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hxa.cpu().numpy(), want_hxa)
    np.testing.assert_array_equal(hms[0].cpu().numpy(), want_mean)
    np.testing.assert_array_equal(hms[1].cpu().numpy(), want_sprd)
    assert np.sum(want_mean != 999.0) > n // 2
