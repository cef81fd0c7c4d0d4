"""Host mirror of the reference's pressure equation, type(poissonEqn)
(src/poissonEqn/poissonEqn.f90:32-63): one channel instance's pressure step
with every field resident on the GPU.

  PoissonEqn(grid, mode, rho_l, rho_g)     poissonEqnCTOR  (:89-150), rho0 = min(rho_l, rho_g) (:127)
  solvePoissonEqn(psi, rho, ux, uy, uz)    solvePoissonEqn (:347-370): source (+ beta), solve
  makeVelocityDivFree(psi, rho, ux, ...)   momentumEqn.f90:795-869 (MG) / 999-1127 (FAST)
  updatePressure(psi, p)                   updatePressure  (:374-459): p = psi, old levels rotated

FAST_MODE keeps two old pressure levels, psi^n and psi^(n-1) (allocateOldField,
poissonEqn.f90:145-147), used by the splitting source and projection; they
start absent (zero), as in a run from rest.  Surface tension is not part of
this path.  Fields are CUDA tensors of the interior cells (shape (nz, ny, nx));
uy has ny+1 faces per column.
"""
from . import fields as F
from .solvers import FastPoissonSolver, PCG


class PoissonEqn:
    def __init__(self, grid, mode, rho_l, rho_g, precision="fp64", **pcg_params):
        if mode not in ("FAST", "MG"):
            raise ValueError("mode must be 'FAST' or 'MG'")
        self.grid = grid
        self.mode = mode
        self.rho0_ = min(float(rho_l), float(rho_g))
        self.s_ = None
        self.beta_ = None
        self.old_ = [None, None]  # psi^n, psi^(n-1) (FAST_MODE)
        if mode == "FAST":
            self.fftSolver_ = FastPoissonSolver(grid, precision=precision)
        else:
            self.pcgs_ = PCG(grid, precision=precision, **pcg_params)

    def _work(self, like):
        import torch

        if self.s_ is None:
            self.s_ = torch.zeros_like(like)
            if self.mode == "MG":
                self.beta_ = torch.zeros_like(like)
        return self.s_

    def solvePoissonEqn(self, psi, rho, ux, uy, uz, alpha_dt, stream=None):
        """FAST: s = splitting source, solveFPS.  MG: s = div u*, beta = alpha dt / rho,
        solvePCG (psi is the warm start)."""
        s = self._work(rho)
        if self.mode == "FAST":
            F.computeSourceFast(self.grid, alpha_dt, self.rho0_, rho, ux, uy, uz, s, self.old_[0], self.old_[1],
                                stream=stream)
            self.fftSolver_.solveFPS(psi, s, stream=stream)
        else:
            F.computeSource(self.grid, ux, uy, uz, s, stream=stream)
            F.computeBeta(self.grid, alpha_dt, rho, self.beta_, stream=stream)
            self.pcgs_.solvePCG(psi, self.beta_, s, stream=stream)
        return psi

    def makeVelocityDivFree(self, psi, rho, ux, uy, uz, alpha_dt, stream=None):
        if self.mode == "FAST":
            F.makeVelocityDivFreeFast(self.grid, alpha_dt, self.rho0_, rho, psi, ux, uy, uz, self.old_[0],
                                      self.old_[1], stream=stream)
        else:
            F.makeVelocityDivFree(self.grid, alpha_dt, rho, psi, ux, uy, uz, stream=stream)

    def updatePressure(self, psi, p):
        """p = psi; psi^(n-1) <- psi^n <- psi (storeOldField)."""
        p.copy_(psi)
        prev = self.old_[1]
        self.old_[1] = self.old_[0]
        if prev is None:
            prev = psi.clone()
        else:
            prev.copy_(psi)
        self.old_[0] = prev

    def close(self):
        if self.mode == "FAST":
            self.fftSolver_.close()
        else:
            self.pcgs_.close()
