"""Binary output in the reference's N-to-N format and restart input from it (host side).

The format and registration order are those of SRC/IO/io_binary.c:19-132 with the
registry of SRC/GRID/grid.c:229-232 and SRC/HYDRO_CORE/hydro_core.c:630-935; the file
is parsed here by an independent reader written from that description."""
import ctypes as C
import os

import numpy as np
import pytest

import fe_py

NH = 3


def read_binary(path):
    """[(name, ndims, extents, float32 data)] in file order."""
    out = []
    with open(path, "rb") as fh:
        raw = fh.read()
    pos = 0
    while pos < len(raw):
        (n,) = np.frombuffer(raw, "<i4", 1, pos); pos += 4
        name = raw[pos:pos + n].decode(); pos += n
        (nd,) = np.frombuffer(raw, "<i4", 1, pos); pos += 4
        ext = tuple(int(e) for e in np.frombuffer(raw, "<i4", nd, pos)); pos += 4 * nd
        cnt = int(np.prod(ext))
        data = np.frombuffer(raw, "<f4", cnt, pos).reshape(ext); pos += 4 * cnt
        out.append((name, int(nd), ext, data))
    return out


def _params(tmp_path, **kw):
    p = dict(fe_py.CFG0, Nx=12, Ny=8, Nz=10, outPath=str(tmp_path) + "/", outFileBase="FE_T", ioOutputMode=1)
    p.update(kw)
    return p


def _write(m, step):
    L = m.L
    L.ioWriteBinaryoutFileSingleTime.argtypes = [C.c_int] * 5
    rc = L.ioWriteBinaryoutFileSingleTime(step, m.geti("Nxp"), m.geti("Nyp"), m.geti("Nzp"), NH)
    assert rc == 0


def test_binary_output_records_and_order(tmp_path):
    """Dry CBL with SGS stress output: names in registration order, 3-D/2-D records with
    halo-inclusive extents, rho and p' raw, u..theta and TKE per unit mass (float division
    by rho), and the host state left unchanged by the write."""
    p = _params(tmp_path / "o", hydroSubGridWrite=1)
    os.makedirs(tmp_path / "o", exist_ok=True)
    m = fe_py.Model(p, str(tmp_path), device=False)
    try:
        fe_py.turbulent_state((m.array("hydroFlds"), m.array("sgstkeScalars")), seed=2)
        h0 = m.array("hydroFlds").copy()
        _write(m, 40)
        path = tmp_path / "o" / "FE_T_rank_0.40"
        recs = read_binary(path)
        names = [r[0] for r in recs]
        assert names == ["xPos", "yPos", "zPos", "topoPos", "rho", "u", "v", "w", "theta", "pressure", "TKE_0",
                         "Tau11", "Tau21", "Tau31", "Tau32", "Tau22", "Tau33", "TauTH1", "TauTH2", "TauTH3",
                         "tskin", "fricVel", "htFlux", "invOblen", "z0m", "z0t"]
        e3 = (12 + 2 * NH, 8 + 2 * NH, 10 + 2 * NH)
        for name, nd, ext, data in recs:
            assert (nd, ext) == ((3, e3) if name not in ("topoPos", "tskin", "fricVel", "htFlux", "invOblen", "z0m", "z0t")
                                 else (2, e3[:2]))
        d = {r[0]: r[3] for r in recs}
        h = m.array("hydroFlds")
        np.testing.assert_array_equal(h, h0)            # host state untouched
        np.testing.assert_array_equal(d["rho"], h[0])
        for i, nm in enumerate(["u", "v", "w", "theta"]):
            np.testing.assert_array_equal(d[nm], (h[i + 1] / h[0]).astype(np.float32))
        np.testing.assert_array_equal(d["TKE_0"], (m.array("sgstkeScalars")[0] / h[0]).astype(np.float32))
        np.testing.assert_array_equal(d["zPos"], m.array("zPos"))
        np.testing.assert_array_equal(d["htFlux"], m.array("htFlux"))
    finally:
        m.close()


def test_binary_output_moist_forcing_names(tmp_path):
    """Moist case with tendency and SGS output: F_* records after their fields, the moist
    surface fields, and the moisture SGS fluxes TauQv1..3, TauQl1..3 last."""
    p = dict(fe_py.CFG2, Nx=8, Ny=8, Nz=8, outPath=str(tmp_path) + "/", outFileBase="FE_M", ioOutputMode=1,
             hydroForcingWrite=1, hydroSubGridWrite=1)
    m = fe_py.Model(p, str(tmp_path), device=False)
    try:
        _write(m, 0)
        names = [r[0] for r in read_binary(tmp_path / "FE_M_rank_0.0")]
        assert names == ["xPos", "yPos", "zPos", "topoPos", "rho", "u", "v", "w", "theta", "pressure",
                         "F_rho", "F_u", "F_v", "F_w", "F_theta", "TKE_0", "F_TKE0",
                         "Tau11", "Tau21", "Tau31", "Tau32", "Tau22", "Tau33", "TauTH1", "TauTH2", "TauTH3",
                         "tskin", "fricVel", "htFlux", "invOblen", "qskin", "qFlux", "z0m", "z0t",
                         "qv", "ql", "F_qv", "F_ql", "TauQv1", "TauQv2", "TauQv3", "TauQl1", "TauQl2", "TauQl3"]
    finally:
        m.close()


def test_restart_reads_state_and_step(tmp_path):
    """A run started with inFile = <written file> takes the state back (rho and 2-D fields
    bit for bit, the rho-weighted fields as (field/rho)*rho, within an ulp of the original)
    and the restart step from the file name (time_integration.c:99-105): simTime_it =
    simTime_itRestart = 40, simTime = 40 dt; the default initial conditions are not applied."""
    out = tmp_path / "o"
    os.makedirs(out, exist_ok=True)
    p = _params(out)
    m = fe_py.Model(p, str(tmp_path / "a"), device=False)
    try:
        fe_py.turbulent_state((m.array("hydroFlds"), m.array("sgstkeScalars")), seed=4)
        m.array("htFlux")[...] = np.float32(0.123)
        m.array("fricVel")[...] = np.float32(0.456)
        h0 = m.array("hydroFlds").copy()
        e0 = m.array("sgstkeScalars").copy()
        _write(m, 40)
    finally:
        m.close()
    m = fe_py.Model(dict(p, inPath=str(out) + "/", inFile="FE_T_rank_0.40"), str(tmp_path / "b"), device=False)
    try:
        assert m.geti("simTime_it") == 40 and m.geti("simTime_itRestart") == 40
        assert m.getf("simTime") == pytest.approx(40 * p["dt"], rel=1e-6)
        h = m.array("hydroFlds")
        np.testing.assert_array_equal(h[0], h0[0])
        for f in range(1, 5):
            np.testing.assert_allclose(h[f], h0[f], rtol=2.5e-7, atol=0)
        np.testing.assert_allclose(m.array("sgstkeScalars"), e0, rtol=2.5e-7, atol=0)
        assert np.all(m.array("htFlux") == np.float32(0.123)) and np.all(m.array("fricVel") == np.float32(0.456))
    finally:
        m.close()


def test_restart_rejects_mismatched_grid(tmp_path):
    out = tmp_path / "o"
    os.makedirs(out, exist_ok=True)
    p = _params(out)
    m = fe_py.Model(p, str(tmp_path / "a"), device=False)
    try:
        _write(m, 8)
    finally:
        m.close()
    with pytest.raises(RuntimeError):
        fe_py.Model(dict(p, Nx=16, inPath=str(out) + "/", inFile="FE_T_rank_0.8"), str(tmp_path / "b"), device=False)
