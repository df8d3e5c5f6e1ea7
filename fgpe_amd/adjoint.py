"""Gradients of real-time GPE evolution through the hand-written adjoint.

`GPEEvolve` is a torch.autograd.Function over nsteps real-time RK4 steps of
the one-component GPE: its forward runs the HIP path with a checkpoint every
`ckpt_every` steps (about sqrt(nsteps)), its backward runs the discrete
adjoint of the RK4 step (fgpe_adjoint_backward) and returns gradients with
respect to the trap V, the contact length a, the strength s of the solver's
k-space kernel (and the initial state).  Gradients follow torch's convention
for complex tensors: dL/dRe + i dL/dIm.

    solver = AdjointSolver(grid, dt, nsteps, ckpt_every, dip=(0, 0, 1))
    psiT = GPEEvolve().apply(V, a, a_dd, psi0, solver)
    loss = fidelity(psiT, target, grid.dv); loss.backward()

The kernel is dipolar (s = a_dd, direction `dip`) unless the solver is made
with softcore=(dim, R): then s is the soft-core strength U of the step
potential of radius R (src/parallel.cu:371-477), and a_dd is 0.

With AdjointSolver(..., ncomp=2) the same call evolves two coupled components
(hamiltonians/BEC_twoComponent.cu:168-244): V and psi0 carry a leading axis
of 2, a is the three contact lengths (a_11, a_22, a_12) and its gradient has
the same shape; the k-space kernel acts on the total density.

In terms of g = 4 pi a and eps_dd = a_dd / a:
    dL/dg = (dL/da + eps_dd dL/da_dd) / (4 pi),  dL/deps_dd = a dL/da_dd.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import FGPE_PREC_F64, Context


class AdjointSolver:
    """One device context reused by every forward/backward of GPEEvolve.

    With a torch module given, the context runs on torch's current stream, so
    device tensors pass in and out without host copies or extra
    synchronisation."""

    def __init__(self, grid, dt, nsteps, ckpt_every=None, dip=(0.0, 0.0, 1.0), precision=FGPE_PREC_F64,
                 torch=None, softcore=None, autotune=True, ncomp=1):
        if ncomp not in (1, 2):
            raise ValueError("ncomp is 1 or 2")
        self.ncomp = ncomp
        self.tune = autotune   # run the plan autotuner once, at the first forward
        self.grid = grid
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.every = int(ckpt_every or max(1, round(math.sqrt(nsteps))))
        self.dip = tuple(dip)
        self.softcore = tuple(softcore) if softcore is not None else None
        self.ctx = Context(grid.nx, grid.ny, grid.nz, precision)
        if torch is not None:
            self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        kx, ky, kz = grid.kaxes()
        self.ctx.set_kspace(kx, ky, kz, grid.dv)

    def _contact(self, a):
        if self.ncomp == 1:
            return (float(a),)
        aa = tuple(float(x) for x in np.asarray(a, dtype=np.float64).reshape(-1))
        if len(aa) != 3:
            raise ValueError("two components take a = (a_11, a_22, a_12)")
        return aa

    def _hamiltonian(self, a, s):
        aa = self._contact(a)
        if self.softcore is None:
            self.ctx.set_hamiltonian(ncomp=self.ncomp, a=aa, a_dd=float(s), dip=self.dip)
        else:
            dim, R = self.softcore
            self.ctx.set_hamiltonian(ncomp=self.ncomp, a=aa, sc_U=float(s), sc_R=float(R), sc_dim=int(dim))

    def _autotune_once(self):
        if self.tune:
            self.ctx.autotune(3)   # leaves the uploaded state as it is
            self.tune = False

    def configure(self, V, a, a_dd, psi0):
        """V, psi0 as host arrays; a_dd: the kernel strength (U for soft-core)."""
        self._hamiltonian(a, a_dd)
        self.ctx.upload_vext(np.asarray(V, dtype=np.float64).astype(np.complex128))
        self.ctx.upload_psi(np.asarray(psi0, dtype=np.complex128))
        self._autotune_once()

    def forward(self, V, a, a_dd, psi0):
        """Host arrays in, psi(T) as a host array."""
        self.configure(V, a, a_dd, psi0)
        self.ctx.adjoint_forward(self.dt, self.nsteps, self.every)
        return self.ctx.download_psi()

    def forward_dev(self, torch, V, a, a_dd, psi0):
        """Device tensors in (V float64, psi0 complex128, contiguous, on the
        context's device; a leading axis of 2 for two components), psi(T) as
        a new device tensor."""
        V = V.detach().to(torch.float64).contiguous()
        psi0 = psi0.detach().to(torch.complex128).contiguous()
        n = self.grid.nx * self.grid.ny * self.grid.nz
        if V.numel() != self.ncomp * n or psi0.numel() != self.ncomp * n:
            raise ValueError("V and psi0 must hold ncomp x nx*ny*nz values")
        Vc, pc = V.reshape(self.ncomp, n), psi0.reshape(self.ncomp, n)
        torch.cuda.synchronize()   # the tensors' producers may run on another stream
        self._hamiltonian(a, a_dd)
        for c in range(self.ncomp):
            self.ctx.upload_vext_real_dev(Vc[c].data_ptr(), c)
            self.ctx.upload_psi_dev(pc[c].data_ptr(), c)
        self._autotune_once()
        self.ctx.adjoint_forward(self.dt, self.nsteps, self.every)
        out = torch.empty_like(psi0)
        oc = out.reshape(self.ncomp, n)
        for c in range(self.ncomp):
            self.ctx.download_psi_dev(oc[c].data_ptr(), c)
        self.ctx.sync()   # V and psi0 may be freed once this returns
        return out

    def backward(self, torch, lamT):
        """lamT: torch complex128 tensor on the device (dL/dpsi(T)).  Returns
        dL/dV, dL/da (a float, or (a_11, a_22, a_12) for two components), the
        kernel strength's gradient and dL/dpsi0."""
        lam = lamT.detach().to(torch.complex128).contiguous()
        lam0 = torch.empty_like(lam)
        gV = torch.empty(lam.shape, dtype=torch.float64, device=lam.device)
        torch.cuda.synchronize()
        g = self.ctx.adjoint_backward(lam.data_ptr(), lam0.data_ptr(), gV.data_ptr())
        self.ctx.sync()
        ga = g[0] if self.ncomp == 1 else (g[0], g[3], g[4])
        return gV, ga, (g[1] if self.softcore is None else g[2]), lam0


def _autograd():
    import torch

    class GPEEvolve(torch.autograd.Function):
        @staticmethod
        def forward(fctx, V, a, a_dd, psi0, solver):
            fctx.solver = solver
            fctx.vdev, fctx.pdev = V.device, psi0.device
            fctx.ashape = tuple(a.shape) if torch.is_tensor(a) else ()
            av = a.detach().cpu().numpy() if torch.is_tensor(a) else a
            return solver.forward_dev(torch, V.to("cuda"), av, float(a_dd), psi0.to("cuda")).to(psi0.device)

        @staticmethod
        def backward(fctx, grad_psiT):
            gV, ga, gadd, lam0 = fctx.solver.backward(torch, grad_psiT.to("cuda"))
            gat = torch.tensor(ga, dtype=torch.float64).reshape(fctx.ashape)
            return (gV.to(fctx.vdev), gat, torch.tensor(gadd, dtype=torch.float64), lam0.to(fctx.pdev), None)

    return GPEEvolve


_fn = None


def GPEEvolve():
    """The autograd.Function class (built on first use so that importing this
    module does not import torch)."""
    global _fn
    if _fn is None:
        _fn = _autograd()
    return _fn


def fidelity(psi, target, dv):
    """|<target|psi>|^2 with <f|g> = sum conj(f) g dv (torch tensors); for
    two components (a leading axis of 2) the sum of the components'."""
    if psi.dim() == 4:
        c = (target.conj() * psi).sum(dim=(1, 2, 3)) * dv
        return (c.conj() * c).real.sum()
    c = (target.conj() * psi).sum() * dv
    return (c.conj() * c).real
