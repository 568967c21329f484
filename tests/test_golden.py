"""The CPU oracle against its committed regression fixtures (tests/golden/*.npz).

The fixtures hold seeded states and the FP64/FP32 oracle tendencies of RK stages
0 and 2 (made by tests/golden/make_golden.py).  Every phase of the oracle is a
per-cell map, so a rebuild must reproduce them bit for bit; a change in any
oracle term shows up here before it can move the GPU parity target.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

import fe_py  # noqa: E402


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_oracle_reproduces_golden(name, tmp_path):
    ref = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    params, seed, moist = make_golden.CASES[name]
    assert json.loads(str(ref["params"])) == json.loads(json.dumps(params))
    m = make_golden.state_of(params, seed, moist, str(tmp_path))
    try:
        for host, _, _, _ in make_golden.FIELDS:
            if host in ref:
                np.testing.assert_array_equal(m.array(host), ref[host])
    finally:
        m.close()
    out = make_golden.run_case(params, seed, moist)
    for k in ref.files:
        if k.endswith("_f64") or k.endswith("_f32"):
            np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
