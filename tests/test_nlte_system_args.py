"""Argument checks of snapi_nlte_response_system that return before any
device work (CPU only): sizes outside the kernels' limits and atoms the
system's layout does not describe are refused with SNAPI_ERR_ARG or
SNAPI_ERR_UNSUPPORTED."""
import numpy as np
import pytest

from snapi_amd import abi, nlteresp, workloads
from snapi_amd._native import SnapiError


@pytest.fixture(scope="module")
def case():
    return workloads.nlte_system_case()


def _call(atom, c, **over):
    a = dict(theta=c["theta"], wray=c["wray"], lam=c["lam"], lamw=c["lamw"], sigma=c["sigma"], depth=c["depth"],
             pops=c["pops"], damp=c["damp"], coll=c["coll"], vr=c["vr"], opem=c["opem"])
    a.update(over)
    return nlteresp.response_system(atom, a["theta"], a["wray"], a["lam"], a["lamw"], a["sigma"], a["depth"],
                                    a["pops"], a["damp"], a["coll"], a["vr"], a["opem"], tau_grid=c["tau_grid"])


def _atom(c, **over):
    """The case's atom rebuilt from its tables with some of them replaced."""
    d = dict(np.load(workloads.os.path.join(workloads.GOLDEN, "nlterin.npz")))
    lev, tr = d["levels"], d["trans"]
    stage = lev[:, 1].astype(int)
    ip = np.zeros(3)
    for i in range(len(stage)):
        ip[stage[i]] = lev[i, 5]
    lower = np.array([int(t[3]) for t in tr])  # stage 1 starts the map
    upper = np.array([int(t[5]) for t in tr])
    a = dict(level_stage=stage, level_energy=lev[:, 3], level_g=lev[:, 4], stage_ip=ip, tr_type=tr[:, 1].astype(int),
             tr_lower=lower, tr_upper=upper, tr_A=tr[:, 6], tr_Blu=tr[:, 7], tr_Bul=tr[:, 8],
             tr_damp_rad=d["damp"][0, :, 0], tr_lambda_crit=tr[:, 9], mass=d["consts"][0])
    a.update(over)
    return abi.make_nlte_atom(**a)


def _status(fn):
    with pytest.raises(SnapiError) as e:
        fn()
    return e.value.code


def test_rebuilt_atom_is_the_case_atom(case):
    atom, c = case
    b = _atom(c)
    for f, _ in abi.SnapiNlteAtom._fields_:
        va, vb = getattr(atom, f), getattr(b, f)
        if isinstance(va, int) or isinstance(va, float):
            assert va == vb, f


def test_sizes_outside_the_limits_are_refused(case):
    atom, c = case
    # more than 16 rays
    th = np.full(17, 0.5)
    rc = _status(lambda: _call(atom, c, theta=th, wray=th, vr=np.zeros((17, c["nd"])),
                                opem=np.ones((17, c["nlam"], 8, c["nd"]))))
    assert rc == abi.SNAPI_ERR_UNSUPPORTED


def test_atoms_outside_the_layout_are_refused(case):
    atom, c = case
    # a line across two stages
    bad_up = np.array([int(x) for x in _atom(c)._keep[4]])
    bad_up[0] = 5
    rc = _status(lambda: _call(_atom(c, tr_upper=bad_up), c))
    assert rc == abi.SNAPI_ERR_UNSUPPORTED
    # levels not grouped by stage: the top stage's level first
    rc = _status(lambda: _call(_atom(c, level_stage=np.array([2, 1, 1, 1, 1, 1])), c))
    assert rc == abi.SNAPI_ERR_UNSUPPORTED
    # two levels in the top stage
    rc = _status(lambda: _call(_atom(c, level_stage=np.array([1, 1, 1, 1, 2, 2])), c))
    assert rc == abi.SNAPI_ERR_UNSUPPORTED
    # a transition index out of the map
    bad_lo = np.array([int(x) for x in _atom(c)._keep[3]])
    bad_lo[2] = 9
    rc = _status(lambda: _call(_atom(c, tr_lower=bad_lo), c))
    assert rc == abi.SNAPI_ERR_ARG
