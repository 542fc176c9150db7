"""Reference column dumps under tests/golden/ as library inputs.  Test
infrastructure."""
import os

import numpy as np

from snapi_amd import workloads

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name))


def cfg0():
    g = workloads.cfg0_column()
    d = load("cfg0_synth.npz")
    g["state"] = g["state"][None]
    g.update(stokes_steps=d["stokes_steps"], opem=d["opem"], spop=d["spop"], sdepth=d["sdepth"], zc=d["zc"])
    return g
