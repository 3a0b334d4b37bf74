"""Optional torch.autograd binding of the equilibrium solve.

``EquilibriumSolve.apply(params, load, ctx, opts)`` returns the equilibrium
displacement u(params, load) of an efem_amd.Context (r(u; params, load) = 0)
as a float64 tensor.  Its backward is the hand-written adjoint: one Krylov
solve K_t(u)^T lambda = dL/du on the true tangent, preconditioned by the
approximate-Jacobian Cholesky factor, then
  dL/dparams = -lambda . dr/dparams,   dL/dload = -lambda . dr/dload
(efem_adjoint in include/efem_gpu.h).  The solve and the adjoint run on the
GPU through the C ABI; torch only carries the tensors.

params: [4] or [n_elems, 4] (the layout of efem_problem.params);
load:   scalar follower pressure, or [3] directed traction.
"""
from __future__ import annotations

import numpy as np
import torch

from . import KRYLOV_AUTO, Context

DEFAULTS = {
    "solver": "newton",        # "newton" (true tangent) or "lbfgs" (the reference loop)
    "epsilon": 1e-9,           # ||r||_inf stop of the forward solve
    "max_iterations": 100,
    "krylov_tol": 1e-10,
    "krylov_max": 2000,
    "krylov_method": KRYLOV_AUTO,
    "warm_start": True,        # start from the previous equilibrium
}


class EquilibriumSolve(torch.autograd.Function):
    @staticmethod
    def forward(fctx, params, load, ectx: Context, opts=None):
        o = dict(DEFAULTS, **(opts or {}))
        p = params.detach().to("cpu", torch.float64).numpy().copy()
        ld = load.detach().to("cpu", torch.float64).numpy().copy()
        ectx.set_params(p)
        directed = ectx.mesh.pdir is not None
        if directed:
            ectx.set_load(pdir=ld.reshape(3))
        elif ectx.mesh.faces.size:
            ectx.set_load(pressure=float(ld.reshape(-1)[0]))
        u0 = getattr(ectx, "_eq_u", None) if o["warm_start"] else None
        if o["solver"] == "newton":
            u, st = ectx.newton(u0, epsilon=o["epsilon"], max_iterations=o["max_iterations"],
                                tol=o["krylov_tol"], max_krylov=o["krylov_max"], method=o["krylov_method"])
        else:
            u, st = ectx.solve(u0, epsilon=o["epsilon"], epsilon_rel=0.0, max_iterations=o["max_iterations"])
        if st["status"] != 0 or not st["rinf"] <= o["epsilon"]:
            raise RuntimeError(f"equilibrium solve did not converge: {st}")
        ectx._eq_u = u
        fctx.ectx = ectx
        fctx.opts = o
        fctx.p = p
        fctx.ld = ld
        fctx.u = u
        fctx.pshape = params.shape
        fctx.lshape = load.shape
        return torch.from_numpy(u.copy()).to(params.device)

    @staticmethod
    def backward(fctx, grad_u):
        ectx: Context = fctx.ectx
        o = fctx.opts
        # the context may have been re-parameterised since forward
        ectx.set_params(fctx.p)
        if ectx.mesh.pdir is not None:
            ectx.set_load(pdir=fctx.ld.reshape(3))
        elif ectx.mesh.faces.size:
            ectx.set_load(pressure=float(fctx.ld.reshape(-1)[0]))
        g = grad_u.detach().to("cpu", torch.float64).numpy()
        _, gp, gl, st = ectx.adjoint(fctx.u, g, tol=o["krylov_tol"], max_iterations=o["krylov_max"],
                                     method=o["krylov_method"])
        if not st["converged"]:
            raise RuntimeError(f"adjoint Krylov solve did not converge: {st}")
        dev = grad_u.device
        gparams = torch.from_numpy(np.asarray(gp).reshape(fctx.pshape)).to(dev)
        n = int(np.prod(fctx.lshape)) if len(fctx.lshape) else 1
        gload = torch.from_numpy(gl[:n].copy()).reshape(fctx.lshape).to(dev)
        return gparams, gload, None, None


def solve_equilibrium(ectx: Context, params: torch.Tensor, load: torch.Tensor, **opts) -> torch.Tensor:
    """u(params, load) with the adjoint as its gradient."""
    return EquilibriumSolve.apply(params, load, ectx, opts)
