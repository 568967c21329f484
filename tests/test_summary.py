"""Pins of the state-summary restatement (oracle/fe_summary.py) on hand-made fields."""
import numpy as np

import fe_summary

NH = 3


def _blank(n=(4, 3, 5)):
    return np.ones(tuple(s + 2 * NH for s in n), dtype=np.float32)


def test_extremes_first_location_and_nonfinite():
    """Ties keep the first cell in i-j-k order; NaN/Inf are counted and skipped; the mean
    divides the finite sum by every interior cell; the integral weights by D_Jac dV."""
    f = _blank() * 2.0
    rho = _blank()
    dj = _blank()
    f[NH + 1, NH + 0, NH + 4] = 7.0
    f[NH + 2, NH + 2, NH + 0] = 7.0          # tie, later in i order
    f[NH + 0, NH + 2, NH + 3] = -1.0
    f[NH + 3, NH + 1, NH + 1] = np.nan
    f[NH + 3, NH + 1, NH + 2] = np.inf
    r = fe_summary.field_entry(f, rho, False, dj, 10.0, NH)
    assert (r["max"], r["imax"], r["jmax"], r["kmax"]) == (7.0, NH + 1, NH + 0, NH + 4)
    assert (r["min"], r["imin"], r["jmin"], r["kmin"]) == (-1.0, NH + 0, NH + 2, NH + 3)
    assert r["nan"] == 1 and r["inf"] == 1
    n = 4 * 3 * 5
    finite = 2.0 * (n - 5) + 7.0 + 7.0 - 1.0
    assert r["mean"] == finite / n
    # the integral sums the raw field (NaN propagates into it, as it would into a budget)
    assert np.isnan(r["integral"])


def test_rho_weighted_fields_use_float_reciprocal():
    """u = (rho u) * float(1/rho), as the reference's float arithmetic: with rho = 1.001 and
    rho u = 0.70070004 that gives 0.70000005, one ulp above the divided value 0.7."""
    rho = _blank() * np.float32(1.001)
    ru = _blank() * np.float32(0.70070004)
    r = fe_summary.field_entry(ru, rho, True, _blank(), 1.0, NH)
    assert r["max"] == float(np.float32(0.70000005)) and r["min"] == r["max"]
    assert r["max"] != float(np.float32(np.float32(0.70070004) / np.float32(1.001)))
    assert abs(r["integral"] - float(np.float32(0.70070004)) * 60) < 1e-12
