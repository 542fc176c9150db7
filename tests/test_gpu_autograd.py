"""torch.autograd wrappers over the tangent and adjoint sweeps: gradients of a
chi^2 of the emergent Stokes vector against the sweeps' own Jacobian
contractions (to rounding) and against central differences of the forward
model (node steps 1e-3 of each value) for the nodes whose Jacobian is the
forward model's derivative."""
import numpy as np
import pytest
import torch

from snapi_amd import abi, api_extra, autograd, invert, workloads

pytestmark = pytest.mark.gpu

NODES = [("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0), ("phi", -2.0)]


@pytest.fixture(scope="module")
def setup():
    t = workloads.template_cfg1()
    rng = np.random.default_rng(5)
    npix = 3
    v = np.zeros((npix, len(NODES)))
    v[:, 0] = rng.uniform(0.3e5, 1.8e5, npix)
    v[:, 1] = rng.normal(0, 1.5e5, npix)
    v[:, 2] = v[:, 1] + rng.normal(0, 0.5e5, npix)
    v[:, 3] = rng.uniform(200, 2500, npix)
    v[:, 4] = v[:, 3] * rng.uniform(0.7, 1.3, npix)
    v[:, 5] = np.arccos(rng.uniform(-0.9, 0.9, npix))
    v[:, 6] = rng.uniform(0.1, 1.4, npix)
    st = np.repeat(t["state"][None], npix, axis=0)
    lam = t["lam_vac"][::6]
    obs = invert.node_jacobian(t["table"], t["ray"], invert.node_model(NODES), st, v * 1.05, lam)[0]
    return dict(t=t, v=v, st=st, lam=lam, obs=obs, model=invert.node_model(NODES))


def test_node_stokes_gradient(setup):
    s = setup
    t = s["t"]
    w = torch.tensor([1.0, 5.0, 5.0, 3.0], dtype=torch.float64)
    obs = torch.from_numpy(s["obs"])
    print("autograd NodeStokes: %d pixels, %d nodes, %d wavelengths" % (len(s["v"]), len(NODES), len(s["lam"])),
          flush=True)

    def loss_of(vals):
        S = autograd.node_stokes(vals, t["table"], t["ray"], s["model"], s["st"], s["lam"])
        return (((S - obs) * w) ** 2).sum()

    x = torch.from_numpy(s["v"].copy()).requires_grad_(True)
    loss = loss_of(x)
    loss.backward()
    g = x.grad.numpy()
    # the same contraction from the Jacobian directly
    S, J = invert.node_jacobian(t["table"], t["ray"], s["model"], s["st"], s["v"], s["lam"])
    r = 2.0 * ((S - s["obs"]) * w.numpy() ** 2).reshape(len(s["v"]), -1)
    gj = np.einsum("pr,prn->pn", r, J)
    assert np.max(np.abs(g - gj) / np.abs(gj).max(axis=0)) < 1e-12
    # central differences of the loss; the inclination node is left out: its
    # Jacobian column carries the reference's rho_V term (cos where -sin
    # belongs, atom_rad_synthesis.cc:999), which the sweeps reproduce and the
    # forward model does not have
    for n in range(len(NODES)):
        if NODES[n][0] == "theta":
            continue
        h = 1e-3 * np.abs(s["v"][:, n])
        vp, vm = s["v"].copy(), s["v"].copy()
        vp[:, n] += h
        vm[:, n] -= h
        for p in range(len(s["v"])):
            lp = float(loss_of(torch.from_numpy(np.where(np.arange(len(s["v"]))[:, None] == p, vp, s["v"]))))
            lm = float(loss_of(torch.from_numpy(np.where(np.arange(len(s["v"]))[:, None] == p, vm, s["v"]))))
            fd = (lp - lm) / (2.0 * h[p])
            err = abs(g[p, n] - fd) / max(abs(fd), 1e-300)
            print("  node %d pixel %d: autograd %.6e differences %.6e (rel %.1e)" % (n, p, g[p, n], fd, err),
                  flush=True)
            assert err < 2e-3


def test_depth_stokes_gradient(setup):
    """The per-depth gradient from one reverse sweep equals the contraction
    of the per-depth responses (snapi_stokes_responses) with dL/dS."""
    s = setup
    t = s["t"]
    st = s["st"][:2]
    ds = np.zeros((2, 2) + st.shape[1:])
    ray = abi.SnapiRay(theta=np.pi, lam_mid=t["ray"].lam_mid, tau_grid=1, hminus=1, hminus_mid=1)
    q0 = np.zeros((2, 7, st.shape[2]))
    for p, (f, sc) in enumerate(autograd._DEPTH_FIELDS):
        q0[:, p] = st[:, f] / sc
    q = torch.from_numpy(q0).requires_grad_(True)
    S = autograd.depth_stokes(q, t["table"], ray, st, ds, s["lam"])
    obs = torch.from_numpy(s["obs"][:2])
    loss = ((S - obs) ** 2).sum()
    loss.backward()
    Sr, R = api_extra.stokes_responses(t["table"], ray, st, ds, s["lam"])  # [ncol][7][nd][nl][4]
    dLdS = 2.0 * (Sr - s["obs"][:2])
    gr = np.einsum("cpzls,cls->cpz", R, dLdS)
    g = q.grad.numpy()
    scale = np.abs(gr).max(axis=2, keepdims=True) + 1e-300
    err = float(np.max(np.abs(g - gr) / scale))
    print("autograd DepthStokes: 2 columns; vs responses contraction max rel %.2e" % err, flush=True)
    assert err < 1e-10
