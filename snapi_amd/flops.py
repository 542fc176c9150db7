"""Algorithmic floating-point operation count of one lambda-point formal solve
(opacity build over ND depths + polarized recurrence), as the synthesis
kernel evaluates it.  add/sub/mul/div/sqrt/exp/exp2/sin = 1, FMA = 2.

The Voigt function dominates and its cost depends on which Humlicek region
each (x, a) falls in, so the count classifies every evaluation of a sample
of the workload (numpy, double) and adds fixed per-depth costs.

Per-evaluation costs (counted from snapi_amd/csrc/rt_common.h):
  setup |v|, s, t, u = (a^2 - v^2, -2av), 1/sqrt(pi) scaling    9
  region 1  t*c / (u + 1/2)                                    + 15
  region 2  t (c1 + c2 u) / (3/4 + u (3 + u))                  + 29
  region 3  two complex Horner chains in t + complex division  + 67
  region 4  degree-6/7 chains in u, exp, sin, sqrt, division   + 109
Fixed costs per depth point: H- 22, Zeeman-line geometry and accumulation
43 per split line and 6 per component, 8 per scalar line, reduction to
(K/eta_I, S) 11, recurrence (optical-depth step, weights, right-hand side,
pivot-free 4x4 solve) 157.
"""
import numpy as np

from . import abi
from .physics import C_LIGHT, K_BOLTZ

REGION_COST = np.array([15, 29, 67, 109]) + 9
H_MINUS = 22
SPLIT_LINE = 43
PER_COMPONENT = 6
SCALAR_LINE = 8
REDUCE = 11
RECURRENCE = 157


def voigt_regions(v, a):
    av = np.abs(v)
    s = av + a
    r = np.full(v.shape, 3, dtype=np.int8)
    r[a > 0.195 * av - 0.176] = 2
    r[s > 5.5] = 1
    r[s >= 15.0] = 0
    return r


def flops_per_lambda_point(table, ray, states, lam):
    """Average operation count per (column, wavelength) over the given
    columns [ncol][nfield][nd] (double) and vacuum wavelengths."""
    states = np.asarray(states, dtype=np.float64)
    ncol, nf, nd = states.shape
    lam = np.asarray(lam, dtype=np.float64)
    total = 0.0
    fixed = nd * (H_MINUS * (1 if ray.hminus else 0) + REDUCE) + (nd - 2) * RECURRENCE
    for il in range(table.nline):
        L = table.line[il]
        T = states[:, abi.F_T]
        vt = states[:, abi.F_VT]
        vl = states[:, abi.F_VLOS]
        bm = states[:, abi.F_BMAG]
        nhi = states[:, abi.F_NHI]
        dld = np.sqrt(2.0 * K_BOLTZ * T / L.mass + vt * vt) * L.lam0 / C_LIGHT
        red = 1.66e-24 / (1.0 / 1.008 + 1.66e-24 / L.mass)
        vm = np.sqrt(8.0 * K_BOLTZ * T / np.pi / red)
        gc = nhi * vm * (vm / 1e6) ** (-L.alpha) * L.sigma
        a = L.lam0 ** 2 / (4 * np.pi * C_LIGHT) * (L.grad + gc) / dld  # [ncol][nd]
        x0 = ((L.lam0 - lam[None, :, None]) - lam[None, :, None] * vl[:, None, :] / C_LIGHT) / dld[:, None, :]
        aa = np.broadcast_to(a[:, None, :], x0.shape)
        if not L.split:
            total += REGION_COST[voigt_regions(x0, aa)].sum() + SCALAR_LINE * x0.size
            continue
        lbd = 4.67e-5 * L.lam0 ** 2 * bm / dld
        c = L.comp0
        for g in range(3):
            for m in range(L.ncomp[g]):
                x = x0 + table.comp[c].shift * lbd[:, None, :]
                total += REGION_COST[voigt_regions(x, aa)].sum() + PER_COMPONENT * x.size
                c += 1
        total += SPLIT_LINE * x0.size
    return total / (ncol * len(lam)) + fixed
