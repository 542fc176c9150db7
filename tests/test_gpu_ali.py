"""ALI on the GPU against the reference's nltepops fixture (Ca II 5 levels +
continuum in the 61-depth column, 444 wavelengths, 2 rays) and against the
oracle on a batch of perturbed columns."""
import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, ali

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def col():
    return _golden.nlte()


def _args(p):
    return (p["atom"], p["grid"], p["colstate"][None], p["vproj"][None], p["gamma"][None], p["bgop"][None],
            p["bgem"][None])


def test_ali_accumulate_fp64_matches_reference(col):
    p = col
    print("ALI accumulate: nd %d, %d wavelengths, %d rays" % (p["nd"], p["nlam"], p["nray"]), flush=True)
    acc = ali.accumulate(*_args(p), p["pop0"][None])[0]
    ref = p["acc"]
    for q, name in enumerate(("Jlu", "Jul", "Lstar", "norm")):
        scale = np.maximum(np.abs(ref[..., q]).max(axis=0), 1e-300)
        err = np.max(np.abs(acc[..., q] - ref[..., q]) / scale)
        print("  %-5s max err %.2e" % (name, err), flush=True)
        assert err < 1e-10, name


def test_ali_newpops_matches_reference(col):
    p = col
    new, rc = ali.newpops(p["atom"], p["colstate"][None], p["coll"][None], p["pop0"][None], p["acc"][None], 1)
    err = np.max(np.abs(new[0] / p["newpop"] - 1.0))
    print("ALI newpops: max rel err %.2e, relchange %.6e (reference %.6e)" % (err, rc[0], p["iters"][0, 1]),
          flush=True)
    # the rate matrix spans tens of orders of magnitude; contracted multiply-adds
    # on the GPU move the solution at the 1e-9 level
    assert err < 1e-8
    assert abs(rc[0] / p["iters"][0, 1] - 1.0) < 1e-8


def test_ali_solve_fp64_matches_reference_nltepops(col):
    p = col
    a = _args(p)
    pops, it, rc = ali.solve(*a, p["coll"][None], p["pop0"][None], tol=1e-2, max_iter=100)
    err = np.max(np.abs(pops[0] / p["ref_final"] - 1.0))
    print("ALI solve FP64: %d iterations (reference %d), max rel err %.2e" % (it[0], len(p["iters"]), err),
          flush=True)
    assert it[0] == len(p["iters"])
    assert err < 1e-6


def test_ali_solve_mixed_against_fp64(col):
    p = col
    a = _args(p)
    p64, it64, _ = ali.solve(*a, p["coll"][None], p["pop0"][None], tol=1e-4, max_iter=400)
    p32, it32, rc32 = ali.solve(*a, p["coll"][None], p["pop0"][None], tol=1e-4, max_iter=400,
                                precision=abi.SNAPI_MIXED)
    err = np.max(np.abs(p32 / p64 - 1.0))
    print("ALI solve to 1e-4: FP64 %d iterations, mixed %d; mixed vs FP64 populations max rel %.2e" % (
        it64[0], it32[0], err), flush=True)
    assert rc32[0] >= 0
    assert err < 1e-3


def test_ali_batch_against_oracle(col):
    """Six columns: the fixture column with temperature-free perturbations of
    the background and the start populations; every column converges on its
    own and matches the oracle's iteration."""
    p = col
    rng = np.random.default_rng(4)
    nc = 6
    cs = np.repeat(p["colstate"][None], nc, 0)
    bgop = p["bgop"][None] * rng.uniform(0.7, 1.3, (nc, 1, 1, 1))
    bgem = p["bgem"][None] * rng.uniform(0.7, 1.3, (nc, 1, 1, 1))
    vp = np.repeat(p["vproj"][None], nc, 0) + rng.normal(0, 2e5, (nc, 1, 1)) * np.array([1.0, -1.0])[None, :, None]
    ga = np.repeat(p["gamma"][None], nc, 0)
    co = np.repeat(p["coll"][None], nc, 0)
    p0 = np.repeat(p["pop0"][None], nc, 0)
    gp, git, grc = ali.solve(p["atom"], p["grid"], cs, vp, ga, bgop, bgem, co, p0, tol=1e-2, max_iter=100)
    q = dict(p)
    q.update(colstate=cs, vproj=vp, gamma=ga, bgop=bgop, bgem=bgem, coll=co)
    op, oit, orc = _oracle.ali_solve(q, p0, 1e-2, 100, ncol=nc)
    err = np.max(np.abs(gp / op - 1.0))
    print("ALI batch: iterations GPU %s oracle %s; max rel err %.2e" % (git.tolist(), oit.tolist(), err), flush=True)
    assert np.array_equal(git, oit)
    assert err < 1e-6
