"""Horizontal-mean profile statistics of a model state (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the checker for the device reduction
`fe_profiles` and the measure on which long runs are compared.  Restates the reference's
post-processing of an output state, tutorials/notebooks/fetutorialfunctions.py:
compute_mean_profiles (:55-83: level means of u, v, theta) and
compute_turb_profiles (:344-431: perturbations from the instantaneous level
means, their covariances, resolved TKE, the mean SGS TKE, and the SGS stress
columns of :406-429 from stress fields emitted at the end of a batch).  Velocities, theta, SGS TKE and
moisture are rho-weighted state divided by rho, as the reference's writers
store them (SRC/IO/io_binary.c:55-97).

Columns (NSTAT = 17), one row per level:
  0 <u>  1 <v>  2 <w>  3 <theta>
  4 <u'u'>  5 <u'w'>  6 <v'v'>  7 <v'w'>  8 <w'w'>  9 resolved TKE 1/2(<u'u'>+<v'v'>+<w'w'>)
  10 <theta'w'>  11 <e> (SGS TKE)  12 <q_v>  13 <q_l>  (moisture in g/kg; 0 when absent)
  14 (<tau13>^2 + <tau23>^2)^1/2  15 <tau33>  16 <tau_theta3>, each averaged from the level
     to the cell centre above it (:426-429; the top level is 0; 0 without stresses)
"""
import numpy as np

NSTAT = 17


def profiles(hydro, tke=None, moist=None, nh=3, tau=None):
    """hydro: [5][Nx+2nh][Ny+2nh][Nz+2nh] rho-weighted state (halo-inclusive, z fastest);
    tke, moist: [n][...] or None; tau: the 9 SGS stress fields in the reference's order
    (11, 21, 31, 32, 22, 33, theta 1..3) or None.  Returns float64 [Nz][NSTAT]."""
    sl = (slice(nh, -nh), slice(nh, -nh), slice(nh, -nh))
    rho = hydro[0][sl].astype(np.float64)
    u, v, w, th = (hydro[f][sl].astype(np.float64) / rho for f in (1, 2, 3, 4))
    mean = lambda a: a.mean(axis=(0, 1))
    nz = rho.shape[2]
    out = np.zeros((nz, NSTAT))
    mu, mv, mw, mt = mean(u), mean(v), mean(w), mean(th)
    up, vp, wp, tp = u - mu, v - mv, w - mw, th - mt
    out[:, 0], out[:, 1], out[:, 2], out[:, 3] = mu, mv, mw, mt
    out[:, 4] = mean(up * up)
    out[:, 5] = mean(up * wp)
    out[:, 6] = mean(vp * vp)
    out[:, 7] = mean(vp * wp)
    out[:, 8] = mean(wp * wp)
    out[:, 9] = mean(0.5 * (up * up + vp * vp + wp * wp))
    out[:, 10] = mean(tp * wp)
    if tke is not None:
        out[:, 11] = mean(tke[0][sl].astype(np.float64) / rho)
    if moist is not None:
        out[:, 12] = mean(moist[0][sl].astype(np.float64) / rho)
        if moist.shape[0] > 1:
            out[:, 13] = mean(moist[1][sl].astype(np.float64) / rho)
    if tau is not None:
        t13, t23, t33, tt3 = (mean(tau[k][sl].astype(np.float64)) for k in (2, 3, 5, 8))
        for c, col in ((14, np.sqrt(t13 ** 2 + t23 ** 2)), (15, t33), (16, tt3)):
            out[:-1, c] = 0.5 * (col[:-1] + col[1:])
    return out
