#!/usr/bin/env python3
"""Write the oracle regression fixtures tests/golden/*.npz.

The reference ships no golden vectors for its time step, so these fixtures are
outputs of this repository's CPU oracle (FP64 and FP32 restatements of the
reference's tendency build) on small seeded states.  They lock the oracle
against accidental change (tests/test_golden.py, CPU) and give the GPU parity
tests a committed target (tests/test_gpu_parity.py).

Each case stores: the .in parameters (JSON), the halo-inclusive host state the
device starts from (float32), and the interior tendencies of RK stages 0 and 2
from the FP64 oracle (float64) and the FP32 oracle (float32).

Run from the repository root after building the library and the oracle:
    python3 tests/golden/make_golden.py [case ...]
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "fasteddy-model_amd"))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import fe_py  # noqa: E402
import oracle_py  # noqa: E402

CASES = {
    "dry_12x10x12": (dict(fe_py.CFG0, Nx=12, Ny=10, Nz=12), 7, False),
    "moist_10x8x10": (dict(fe_py.CFG2, Nx=10, Ny=8, Nz=10, d_zeta=120.0), 8, True),
    "smagorinsky_12x10x12": (dict(fe_py.CFG0, Nx=12, Ny=10, Nz=12, TKESelector=0), 9, False),
}
FIELDS = (("hydroFlds", "HYDRO", "hydroFldsFrhs", "HYDRO_F"), ("sgstkeScalars", "TKE", "sgstkeScalarsFrhs", "TKE_F"),
          ("moistScalars", "MOIST", "moistScalarsFrhs", "MOIST_F"))


def state_of(params, seed, moist, workdir):
    m = fe_py.Model(params, workdir, device=False)
    fe_py.turbulent_state((m.array("hydroFlds"), m.array("sgstkeScalars")), seed=seed)
    if moist:
        fe_py.moist_state(m, seed=seed + 1)
    return m


def run_case(params, seed, moist):
    with tempfile.TemporaryDirectory() as td:
        m = state_of(params, seed, moist, td)
        out = {"params": np.array(json.dumps(params))}
        for host, _, _, _ in FIELDS:
            a = m.array(host)
            if a is not None:
                out[host] = a.copy()
        for prec in (64, 32):
            o = oracle_py.Oracle(params, prec)
            o.load_from_model(m)
            for stage in (0, 2):
                o.build_frhs(stage)
                for host, _, fhost, oname in FIELDS:
                    a = m.array(host)
                    if a is None:
                        continue
                    F = m.interior(o.arr(oname)[: a.shape[0]]).copy()
                    out[f"{fhost}_s{stage}_f{prec}"] = F.astype(np.float64 if prec == 64 else np.float32)
            o.close()
        m.close()
    return out


def main():
    names = sys.argv[1:] or list(CASES)            # regenerate the named cases only, if any given
    for name in names:
        params, seed, moist = CASES[name]
        out = run_case(params, seed, moist)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
