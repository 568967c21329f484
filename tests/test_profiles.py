"""Pins of the profile-statistics restatement (oracle/fe_profiles.py) against closed forms.

Periodic sinusoids sampled at cell centres over whole periods have exactly known
level means and covariances, so the restatement's centring, products and
averaging are checked independently of how it is written."""
import math

import numpy as np

import fe_profiles

NH = 3


def test_profile_statistics_of_sinusoids():
    nx, ny, nz = 24, 16, 5
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    U0, V0, A, B, Cc, D, E = 7.0, -2.0, 1.3, 0.6, 0.9, 0.4, 0.25
    sx, cy = np.sin(2 * math.pi * X), np.cos(2 * math.pi * Y)
    rho = 1.17
    h = np.zeros((5, nx + 2 * NH, ny + 2 * NH, nz + 2 * NH))
    tke = np.zeros((1,) + h.shape[1:])
    moist = np.zeros((2,) + h.shape[1:])
    sl = (slice(NH, -NH), slice(NH, -NH))
    for k in range(NH, NH + nz):
        h[0][sl + (k,)] = rho
        h[1][sl + (k,)] = rho * (U0 + A * sx)
        h[2][sl + (k,)] = rho * (V0 + Cc * cy)
        h[3][sl + (k,)] = rho * (B * sx)
        h[4][sl + (k,)] = rho * (300.0 + D * sx)
        tke[0][sl + (k,)] = rho * E
        moist[0][sl + (k,)] = rho * 12.0
        moist[1][sl + (k,)] = rho * 0.5
    # SGS stresses: level means tau13 = 3k, tau23 = 4k (magnitude 5k), tau33 = k, tau_theta3 = -k
    # with zero-mean fluctuations; on cell centres the columns are 5(k+1/2), k+1/2, -(k+1/2)
    tau = np.zeros((9,) + h.shape[1:])
    for k in range(nz):
        kk = NH + k
        tau[2][sl + (kk,)] = 3.0 * k + sx
        tau[3][sl + (kk,)] = 4.0 * k - cy
        tau[5][sl + (kk,)] = 1.0 * k + sx * cy
        tau[8][sl + (kk,)] = -1.0 * k
    p = fe_profiles.profiles(h, tke, moist, nh=NH, tau=tau)
    expect = [U0, V0, 0.0, 300.0, A * A / 2, A * B / 2, Cc * Cc / 2, 0.0, B * B / 2,
              (A * A + B * B + Cc * Cc) / 4, D * B / 2, E, 12.0, 0.5]
    assert p.shape == (nz, fe_profiles.NSTAT)
    np.testing.assert_allclose(p[:, :14], np.tile(expect, (nz, 1)), rtol=1e-12, atol=1e-12)
    kc = np.arange(nz - 1) + 0.5
    np.testing.assert_allclose(p[:-1, 14], 5.0 * kc, rtol=1e-12)
    np.testing.assert_allclose(p[:-1, 15], kc, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(p[:-1, 16], -kc, rtol=1e-12)
    assert np.all(p[-1, 14:] == 0.0)
