"""Stage tendencies at the production sizes against the FP32 and FP64 oracles on the same state:
the dry CBL (512x512x128 cells, dx = dz = 10 m, dt = 0.02 s) and the moist shallow-cumulus case
(256x256x160 cells, dx = dz = 25 m, dt = 0.05 s, q_v and q_l).

The state is the seeded initial state with smooth random velocity, theta and SGS TKE
perturbations on top (and seeded moisture for the moist case), so that every SGS term is
active.  Per prognostic field the test prints the maximum and RMS error of the GPU tendency
relative to the FP64 oracle's field maximum, next to the FP32 oracle's (the reference's
arithmetic), and bounds the maximum as the small-grid parity tests do:
    max|F_gpu - F_64| <= max(1e-5 * max|F_64|, 2 * max|F_ref32 - F_64|).
"""
import os

import numpy as np
import pytest

import fe_py
import oracle_py

pytestmark = pytest.mark.gpu

CASES = [
    ("512x512x128 dry", dict(fe_py.CFG1, NtBatch=1), False),
    ("256x256x160 moist", dict(fe_py.CFG2, NtBatch=1), True),
]


def _errors(g, r32, r64):
    ax = tuple(range(1, g.ndim))
    scale = np.maximum(np.abs(r64).max(axis=ax), 1e-300)
    e_max = np.abs(g - r64).max(axis=ax) / scale
    e_rms = np.sqrt(((g - r64) ** 2).mean(axis=ax)) / scale
    e_ref = np.abs(r32 - r64).max(axis=ax) / scale
    return e_max, e_rms, e_ref


@pytest.mark.parametrize("label,p,moist", CASES, ids=[c[0] for c in CASES])
def test_production_size_stage_tendencies(label, p, moist, tmp_path):
    threads = min(16, os.cpu_count() or 1)
    m = fe_py.Model(p, str(tmp_path))
    try:
        fe_py.turbulent_state((m.array("hydroFlds"), m.array("sgstkeScalars")), seed=11)
        if moist:
            fe_py.moist_state(m, seed=12)
        m.upload()
        names = ["rho", "rho u", "rho v", "rho w", "rho theta", "rho e"] + (["rho qv", "rho ql"] if moist else [])
        host = [("hydroFldsFrhs", "HYDRO_F", 5), ("sgstkeScalarsFrhs", "TKE_F", 1)]
        if moist:
            host.append(("moistScalarsFrhs", "MOIST_F", 2))
        for stage in (0, 2):
            m.build_frhs(stage)
            g = np.concatenate([m.interior(m.array(h)[:n]) for h, _, n in host]).astype(np.float64)
            ref = {}
            for bits in (32, 64):
                o = oracle_py.Oracle(p, bits, threads=threads)
                try:
                    o.load_from_model(m)
                    o.build_frhs(stage)
                    ref[bits] = np.concatenate([m.interior(o.arr(a)[:n]) for _, a, n in host]).astype(np.float64)
                finally:
                    o.close()
            e_max, e_rms, e_ref = _errors(g, ref[32], ref[64])
            for f, n in enumerate(names):
                print(f"{label} stage {stage} {n:9s}: gpu max {e_max[f]:.2e} rms {e_rms[f]:.2e} | "
                      f"ref32 max {e_ref[f]:.2e}", flush=True)
            assert np.all(np.isfinite(g))
            assert np.all(e_max <= np.maximum(1e-5, 2 * e_ref)), (stage, e_max, e_ref)
            del g, ref
    finally:
        m.close()
