"""Small host-side physics helpers of the reference used to assemble column
state (FP64, numpy).

planck     cmlib/mathtools.cc:1150-1152
airtovac   cmlib/mathtools.cc:1325-1331
THOMSON    electron scattering cross-section of atmos/atmos_rad_synthesis.cc:41
"""
import numpy as np

C_LIGHT = 2.99792458E10
H_PLANCK = 6.6260687652E-27
K_BOLTZ = 1.3806503E-16
THOMSON = 6.65E-25


def planck(lam, T):
    return 2.0 * 6.62606957 * 8.9875517E-7 / lam ** 5 / (np.exp(1.4387769 / lam / T) - 1.0)


def airtovac(lam_air):
    s = 1e4 / (np.asarray(lam_air) * 1e8)
    n = 1.0 + 0.00008336624212083 + 0.02408926869968 / (130.1065924522 - s * s) \
        + 0.0001599740894897 / (38.92568793293 - s * s)
    return lam_air * n
