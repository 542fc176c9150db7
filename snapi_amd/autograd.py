"""torch.autograd.Function wrappers over the C ABI's derivative sweeps, for
Python tooling: gradients of any scalar of the emergent Stokes vector.

NodeStokes: Stokes of node values (snapi_node_jacobian; the forward-mode
  sweep carries every node direction, so the forward pass keeps the Jacobian
  and the backward pass is J^T g).  The node model and the state are as
  snapi_node_jacobian takes them (kinematic and magnetic nodes, mu = 1 ray).
DepthStokes: Stokes of the seven per-depth parameters (T, Nt, vt, vLOS, |B|,
  inclination and azimuth in radians) of a column state; the backward pass is
  one reverse sweep (snapi_stokes_adjoint with the incoming gradient as the
  residual: w = 1 and S_obs = S - g, so sum_s w (S - S_obs) dS/dq = g dS/dq).
  The thermodynamic fields that depend on T and Nt enter the gradient through
  dstate, the reference's response path (obs_stokes_responses,
  atmos/atmos_rad_synthesis.cc:166-399); the forward pass synthesises the
  state as given.

Both run the library's host-pointer entry points: tensors are moved to host
memory for the call and the results come back on the input's device."""
import numpy as np
import torch

from . import abi, api, api_extra, invert

# state field of each per-depth parameter and its scale from the
# parameter's unit to the state's (angles: radians -> degrees)
_DEPTH_FIELDS = [(abi.F_T, 1.0), (abi.F_NT, 1.0), (abi.F_VT, 1.0), (abi.F_VLOS, 1.0), (abi.F_BMAG, 1.0),
                 (abi.F_BINC, 180.0 / np.pi), (abi.F_BAZI, 180.0 / np.pi)]


class NodeStokes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nodes, table, ray, model, state, lam_vac, precision=abi.SNAPI_FP64,
                jacobian=abi.SNAPI_JAC_TANGENT):
        nv = nodes.detach().to("cpu", torch.float64).numpy()
        S, J = invert.node_jacobian(table, ray, model, state, nv, lam_vac, precision=precision, jacobian=jacobian)
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(J, dtype=np.float64)))
        ctx.out_device = nodes.device
        ctx.out_dtype = nodes.dtype
        return torch.from_numpy(np.asarray(S, dtype=np.float64)).to(nodes.device, nodes.dtype)

    @staticmethod
    def backward(ctx, g):
        (J,) = ctx.saved_tensors  # [npix][nlambda*4][nnode]
        gs = g.detach().to("cpu", torch.float64).reshape(J.shape[0], -1)
        gn = torch.einsum("pr,prn->pn", gs, J).to(ctx.out_device, ctx.out_dtype)
        return gn, None, None, None, None, None, None, None


class DepthStokes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, table, ray, state, dstate, lam_vac, precision=abi.SNAPI_FP64):
        q = params.detach().to("cpu", torch.float64).numpy()
        st = np.array(state, dtype=np.float64, copy=True)
        if st.ndim == 2:
            st = st[None]
        if q.shape != (st.shape[0], 7, st.shape[2]):
            raise ValueError("params must be [ncol][7][nd] = %s, got %s" % ((st.shape[0], 7, st.shape[2]), q.shape))
        for p, (f, sc) in enumerate(_DEPTH_FIELDS):
            st[:, f] = q[:, p] * sc
        S = api.synth_stokes(table, ray, st, lam_vac, precision)
        ctx.args = (table, ray, st, np.asarray(dstate), lam_vac, precision)
        ctx.S = np.asarray(S, dtype=np.float64)
        ctx.out_device = params.device
        ctx.out_dtype = params.dtype
        return torch.from_numpy(ctx.S.copy()).to(params.device, params.dtype)

    @staticmethod
    def backward(ctx, g):
        table, ray, st, ds, lam, precision = ctx.args
        gs = g.detach().to("cpu", torch.float64).numpy().reshape(ctx.S.shape)
        grad = api_extra.stokes_adjoint(table, ray, st, ds, lam, ctx.S - gs, (1.0, 1.0, 1.0, 1.0), precision)
        return (torch.from_numpy(np.asarray(grad, dtype=np.float64)).to(ctx.out_device, ctx.out_dtype),
                None, None, None, None, None, None)


def node_stokes(nodes, table, ray, model, state, lam_vac, precision=abi.SNAPI_FP64, jacobian=abi.SNAPI_JAC_TANGENT):
    return NodeStokes.apply(nodes, table, ray, model, state, lam_vac, precision, jacobian)


def depth_stokes(params, table, ray, state, dstate, lam_vac, precision=abi.SNAPI_FP64):
    return DepthStokes.apply(params, table, ray, state, dstate, lam_vac, precision)
