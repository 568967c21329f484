"""Terrain input of the host GRID module (SRC/GRID/grid.c:299-410, metrics :584-741).

A topoFile is raw binary: int Nx, int Ny, then Nx*Ny floats with x fastest.  Each column is
compressed terrain-following, z = zeta (z_top - h) / z_top + h, and the metric tensor comes
from centred differences of the coordinates, T[i][j] = dx_j/dxi_i.  The reference's metric
entries are cofactors of T over its determinant: D = det T, J33 = dzeta/dz, and
J31 = (T10 T21 - T11 T20)/D = dxi/dz, J32 = deta/dz (grid.c:634-636), which vanish whenever x
and y do not vary along a column -- so a terrain-following grid has J31 = J32 = 0 even over a
slope, and the terrain enters only through D and J33.  For a constant elevation h and for a
ramp h = s x the differences are exact away from the periodic seam and the ground/lid halos:
    D = dz/dzeta = (z_top - h) / z_top,  J33 = 1 / D,  J31 = J32 = 0.
(A slope term dzeta/dx = -s (z_top - zeta)/(z_top - h) would be the other cofactor; the test
fails if it appears in J31.)
"""
import numpy as np
import pytest

import fe_py

NH = 3


def write_topo(path, h):
    """h indexed [j][i] (x fastest in the file)."""
    ny, nx = h.shape
    with open(path, "wb") as fh:
        np.array([nx, ny], dtype=np.int32).tofile(fh)
        np.asarray(h, dtype=np.float32).tofile(fh)
    return path


def _model(tmp_path, h, **kw):
    ny, nx = h.shape
    p = dict(fe_py.CFG0, Nx=nx, Ny=ny, Nz=16, d_xi=20.0, d_eta=20.0, d_zeta=20.0,
             topoFile=write_topo(str(tmp_path / "topo.bin"), h), **kw)
    return fe_py.Model(p, str(tmp_path), device=False), p


def _interior(a, nz):
    return a[NH + 1:-NH - 1, NH + 1:-NH - 1, NH + 1:NH + nz - 1]


def test_constant_elevation_metrics(tmp_path):
    h0 = 150.0
    m, p = _model(tmp_path, np.full((12, 16), h0))
    try:
        nz = p["Nz"]
        ztop = nz * p["d_zeta"] - 0.5 * p["d_zeta"]
        zeta = (np.arange(nz) + 0.5) * p["d_zeta"]
        z = m.array("zPos")[NH:-NH, NH:-NH, NH:NH + nz]
        np.testing.assert_allclose(z, np.broadcast_to(zeta * (ztop - h0) / ztop + h0, z.shape), rtol=1e-6)
        np.testing.assert_allclose(m.array("topoPos")[NH:-NH, NH:-NH], h0)
        d = (ztop - h0) / ztop
        np.testing.assert_allclose(_interior(m.array("D_Jac"), nz), d, rtol=1e-5)
        np.testing.assert_allclose(_interior(m.array("J33"), nz), 1.0 / d, rtol=1e-5)
        assert np.abs(_interior(m.array("J31"), nz)).max() < 1e-6
        assert np.abs(_interior(m.array("J32"), nz)).max() < 1e-6
    finally:
        m.close()


def test_ramp_metrics(tmp_path):
    nx, ny, s = 24, 12, 0.05
    x = (np.arange(nx) + 0.5) * 20.0
    h = np.broadcast_to(s * x, (ny, nx)).copy()
    m, p = _model(tmp_path, h)
    try:
        nz = p["Nz"]
        ztop = nz * p["d_zeta"] - 0.5 * p["d_zeta"]
        topo = m.array("topoPos")[NH:-NH, NH:-NH]
        np.testing.assert_allclose(topo, h.T, rtol=1e-6)                   # host arrays are [i][j]
        # halo columns take the nearest interior elevation
        np.testing.assert_allclose(m.array("topoPos")[:NH, NH:-NH], np.broadcast_to(h.T[0], (NH, ny)))
        hh = h.T[:, :, None]
        d = ((ztop - hh) / ztop * np.ones(nz))[1:-1, 1:-1, 1:-1]
        np.testing.assert_allclose(_interior(m.array("D_Jac"), nz), d, rtol=2e-5)
        np.testing.assert_allclose(_interior(m.array("J33"), nz), 1.0 / d, rtol=2e-5)
        assert np.abs(_interior(m.array("J31"), nz)).max() < 1e-6
        assert np.abs(_interior(m.array("J32"), nz)).max() < 1e-6
    finally:
        m.close()


def test_topofile_extent_mismatch_is_an_error(tmp_path):
    path = write_topo(str(tmp_path / "topo.bin"), np.zeros((10, 10)))
    p = dict(fe_py.CFG0, Nx=16, Ny=12, Nz=16, topoFile=path)
    with pytest.raises(RuntimeError):
        fe_py.Model(p, str(tmp_path), device=False)


@pytest.mark.gpu
def test_constant_elevation_steps_match_oracle_gpu(tmp_path):
    """A topoFile of constant elevation (150 m) keeps the metrics horizontally uniform, so the
    device runs it: three RK3 steps against the FP64 oracle, state within 2e-3 of its change."""
    import oracle_py
    h = np.full((16, 64), 150.0)
    p = dict(fe_py.CFG0, Nx=64, Ny=16, Nz=20, NtBatch=1, topoFile=write_topo(str(tmp_path / "topo.bin"), h))
    m = fe_py.Model(p, str(tmp_path / "run"))
    try:
        fe_py.turbulent_state((m.array("hydroFlds"), m.array("sgstkeScalars")), seed=21)
        o64 = oracle_py.Oracle(p, 64)
        o64.load_from_model(m)
        start = {k: m.interior(m.array(k)).astype(np.float64).copy() for k in ("hydroFlds", "sgstkeScalars")}
        m.L.fe_set_int(b"simTime_itRestart", 0)
        for _ in range(3):
            m.commence()
        o64.steps(3)
        for hname, oname in (("hydroFlds", "HYDRO"), ("sgstkeScalars", "TKE")):
            g = m.interior(m.array(hname)).astype(np.float64)
            o = m.interior(o64.arr(oname)[: g.shape[0]])
            change = np.abs(o - start[hname]).max(axis=(1, 2, 3))
            diff = np.abs(g - o).max(axis=(1, 2, 3))
            assert np.all(diff <= 2e-3 * change + 1e-6 * np.abs(o).max(axis=(1, 2, 3))), (hname, diff, change)
        o64.close()
    finally:
        m.close()


@pytest.mark.gpu
def test_sloped_terrain_is_rejected_on_the_device(tmp_path):
    """Horizontally varying metrics (a ramp) are refused at cuda_gridInit with FE_ERR_TERRAIN (105)
    before any kernel runs; the host set-up itself succeeds."""
    x = (np.arange(64) + 0.5) * 20.0
    h = np.broadcast_to(0.05 * x, (16, 64)).copy()
    p = dict(fe_py.CFG0, Nx=64, Ny=16, Nz=20, topoFile=write_topo(str(tmp_path / "topo.bin"), h))
    with pytest.raises(RuntimeError, match="fe_setup_device failed with 105"):
        fe_py.Model(p, str(tmp_path / "run"))
