"""Per-field state summary of a model state (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the device reduction
`fe_state_summary`.  Restates the reference's host scan
hydro_coreFldStateLogEntry (SRC/HYDRO_CORE/hydro_core.c:1422-1525) as called by
hydro_coreStateLogDump (:1282-1416) for rho, u, v, w, theta, the pressure, the
SGS TKE and the moisture species:

  * over the interior cells in the reference's loop order (i outermost, then j,
    then k), the value is the field itself (rho, pressure) or field * rhoInv with
    rhoInv = float(1.0/rho) (:1452-1456: a double quotient stored in a float, then
    a float product);
  * NaN and Inf values are counted and left out of the extremes and the sum;
  * max and min keep their first location (strict comparisons, :1468-1479);
  * mean = sum of the finite values / number of interior cells, all cells counted
    (:1484-1490).  The reference accumulates the sum in float; this restatement
    and the device use FP64, which is the intended difference;
  * integral = sum of the raw field * D_Jac * dX dY dZ (the volume integral: mass,
    rho*theta, ...), a budget column the reference does not print.
"""
import numpy as np

COLS = ("max", "imax", "jmax", "kmax", "min", "imin", "jmin", "kmin", "mean", "nan", "inf", "integral")


def field_entry(fld, rho, cons, djac, dV, nh=3):
    """Summary of one halo-inclusive [Nx+2nh][Ny+2nh][Nz+2nh] float32 field."""
    sl = (slice(nh, -nh), slice(nh, -nh), slice(nh, -nh))
    f = fld[sl].astype(np.float32)
    if cons:
        rinv = (1.0 / rho[sl].astype(np.float64)).astype(np.float32)
        v = (f * rinv).astype(np.float32)
    else:
        v = f
    flat = v.reshape(-1)
    nan = int(np.isnan(flat).sum())
    inf = int(np.isinf(flat).sum())
    ok = np.isfinite(flat)
    out = dict(nan=nan, inf=inf)
    shp = v.shape
    if ok.any():
        vmax = np.where(ok, flat, -np.inf)
        vmin = np.where(ok, flat, np.inf)
        imax, imin = int(np.argmax(vmax)), int(np.argmin(vmin))
        out["max"], out["min"] = float(flat[imax]), float(flat[imin])
        for nm, q in (("max", imax), ("min", imin)):
            i, j, k = np.unravel_index(q, shp)
            out["i" + nm], out["j" + nm], out["k" + nm] = i + nh, j + nh, k + nh
    else:
        out.update(max=-1.0e16, min=1.0e16, imax=0, jmax=0, kmax=0, imin=0, jmin=0, kmin=0)
    out["mean"] = float(flat[ok].astype(np.float64).sum()) / flat.size
    out["integral"] = float((f.astype(np.float64) * djac[sl].astype(np.float64)).sum()) * dV
    return out


def state_summary(hydro, pres, djac, dV, tke=None, moist=None, nh=3):
    """{row: {column: value}} for rho, u, v, w, theta, press, TKE_0, qv, ql."""
    rho = hydro[0]
    rows = {}
    for name, fld, cons in (("rho", hydro[0], False), ("u", hydro[1], True), ("v", hydro[2], True),
                            ("w", hydro[3], True), ("theta", hydro[4], True), ("press", pres, False)):
        rows[name] = field_entry(fld, rho, cons, djac, dV, nh)
    if tke is not None:
        for n in range(tke.shape[0]):
            rows[f"TKE_{n}"] = field_entry(tke[n], rho, True, djac, dV, nh)
    if moist is not None:
        for n, nm in zip(range(moist.shape[0]), ("qv", "ql", "qr")):
            rows[nm] = field_entry(moist[n], rho, True, djac, dV, nh)
    return rows
