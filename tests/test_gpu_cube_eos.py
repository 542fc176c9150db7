"""The synthesis cube workload with its thermodynamics solved on the GPU
(workloads.cube_state_eos_dev: hydrostatic equilibrium, the float32 model
cube, the cube to column states, the equation of state) against the
oracle's restatement of the same chain on the same draws."""
import numpy as np
import pytest

import _oracle
from snapi_amd import abi, workloads

pytestmark = pytest.mark.gpu


def test_eos_cube_matches_oracle_chain():
    t = workloads.template_cfg1()
    nx, ny = 8, 8
    print("cube %d x %d x 71 through hydrostatic equilibrium and the equation of state" % (nx, ny), flush=True)
    st, c32 = workloads.cube_state_eos_dev(t, nx, ny, seed=7)
    got = st.cpu().numpy()
    assert tuple(c32.shape) == (12, nx, ny, 71)
    want = _oracle.cube_state_eos(t, nx, ny, 7, np.arange(nx * ny))
    nf = abi.nfield(t["table"].nline)
    fields = [abi.F_T, abi.F_VT, abi.F_VLOS, abi.F_BMAG, abi.F_BINC, abi.F_BAZI, abi.F_NE, abi.F_NHI, abi.F_OPREF,
              abi.F_GRID, abi.F_BGOP, abi.F_BGEM, abi.F_NT] + list(range(abi.NF_BASE, nf))
    worst = 0.0
    for f in fields:
        e = float(np.max(np.abs(got[:, f] - want[:, f]) / np.maximum(np.abs(want[:, f]), 1e-300)))
        worst = max(worst, e)
        # the float32 model cube can round a pressure one unit apart: 6e-8
        assert e < 1e-6, (f, e)
    print("  all fields within %.1e of the oracle" % worst, flush=True)
    # hydrostatic: the gas pressure rises monotonically with depth in every column
    pg = c32[3].cpu().numpy().reshape(nx * ny, 71)
    assert np.all(np.diff(pg, axis=1) > 0)
