"""Restart from the binary output on the GPU: a run written at step 6 and restarted from
that file reaches step 10 in the state of the uninterrupted run, to the rounding that the
per-unit-mass format and the surface-layer re-initialisation at a restart step
(cuda_surfaceLayerDevice.cu:147-153: c_d and c_h from u* and the heat flux) introduce.
The stand-alone driver writes the reference's output sequence and restarts from it."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import fe_py
from test_io import read_binary

pytestmark = pytest.mark.gpu
NH = 3
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(m, step):
    m.L.ioWriteBinaryoutFileSingleTime.argtypes = [C.c_int] * 5
    assert m.L.ioWriteBinaryoutFileSingleTime(step, m.geti("Nxp"), m.geti("Nyp"), m.geti("Nzp"), NH) == 0


def test_restart_continues_the_run(tmp_path):
    out = tmp_path / "o"
    os.makedirs(out)
    p = dict(fe_py.CFG0, Nx=64, Ny=32, Nz=24, NtBatch=1, outPath=str(out) + "/", outFileBase="FE_R", ioOutputMode=1)
    m = fe_py.Model(p, str(tmp_path / "a"))
    try:
        fe_py.turbulent_state((m.array("hydroFlds"), m.array("sgstkeScalars")), seed=11)
        for _ in range(6):
            m.commence()
        assert m.geti("simTime_it") == 6
        _write(m, 6)
        h6 = m.array("hydroFlds").copy()
        for _ in range(4):
            m.commence()
        hA = m.array("hydroFlds").copy()
        eA = m.array("sgstkeScalars").copy()
    finally:
        m.close()
    m = fe_py.Model(dict(p, inPath=str(out) + "/", inFile="FE_R_rank_0.6"), str(tmp_path / "b"))
    try:
        assert m.geti("simTime_it") == 6
        for _ in range(4):
            m.commence()
        assert m.geti("simTime_it") == 10
        hB = m.array("hydroFlds").copy()
        eB = m.array("sgstkeScalars").copy()
    finally:
        m.close()
    sl = (slice(None), slice(NH, -NH), slice(NH, -NH), slice(NH, -NH))
    change = np.abs(hA[sl] - h6[sl]).max(axis=(1, 2, 3))
    diff = np.abs(hA[sl] - hB[sl]).max(axis=(1, 2, 3))
    scale = np.abs(hA[sl]).max(axis=(1, 2, 3))
    print(f"restart: |A-B| {diff} vs 4-step change {change}")
    assert np.all(diff <= 1e-3 * change + 1e-6 * scale), (diff, change)
    assert np.abs(eA[sl] - eB[sl]).max() <= 1e-3 * np.abs(eA[sl]).max()


def test_driver_output_and_restart(tmp_path):
    """fasteddy_amd case.in: state summaries and binary files at steps 0, 5, 10 (every
    frqOutput = 5 steps and at the end); a second run from FE_D_rank_0.5 ends in the same
    step-10 state to the restart rounding."""
    exe = os.path.join(REPO, "fasteddy-model_amd", "fasteddy_amd")
    assert os.path.exists(exe), "build the stand-alone driver first"
    out = tmp_path / "o"
    os.makedirs(out)
    p = dict(fe_py.CFG0, Nx=64, Ny=32, Nz=24, Nt=10, NtBatch=5, frqOutput=5, outPath=str(out) + "/",
             outFileBase="FE_D", ioOutputMode=1)
    case = fe_py.write_in(p, str(tmp_path / "case.in"))
    r = subprocess.run([exe, case], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "STATE-SUMMARY @ it = 0" in r.stdout and "STATE-SUMMARY @ it = 10" in r.stdout
    assert "CORRUPTED" not in r.stdout
    for s in (0, 5, 10):
        assert (out / f"FE_D_rank_0.{s}").exists()
    a10 = {n: d for n, _, _, d in read_binary(out / "FE_D_rank_0.10")}
    os.rename(out / "FE_D_rank_0.10", out / "first_run.10")
    case2 = fe_py.write_in(dict(p, inPath=str(out) + "/", inFile="FE_D_rank_0.5"), str(tmp_path / "case2.in"))
    r = subprocess.run([exe, case2], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "STATE-SUMMARY @ it = 5" in r.stdout and "STATE-SUMMARY @ it = 10" in r.stdout
    b10 = {n: d for n, _, _, d in read_binary(out / "FE_D_rank_0.10")}
    a5 = {n: d for n, _, _, d in read_binary(out / "FE_D_rank_0.5")}
    sl = (slice(NH, -NH),) * 3
    for n in ("rho", "u", "v", "w", "theta", "TKE_0"):
        a, b, c = (x[n][sl].astype(np.float64) for x in (a10, b10, a5))
        diff, change = np.abs(a - b).max(), np.abs(a - c).max()
        print(f"driver restart {n}: |A-B| {diff:.3e} vs 5-step change {change:.3e}")
        assert diff <= 1e-3 * change + 1e-6 * np.abs(a).max(), n
