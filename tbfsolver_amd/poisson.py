"""Host mirror of the reference's pressure equation, type(poissonEqn)
(src/poissonEqn/poissonEqn.f90:32-63): one channel instance's pressure step
with every field resident on the GPU.

  PoissonEqn(grid, mode, rho_l, rho_g)     poissonEqnCTOR  (:89-150), rho0 = min(rho_l, rho_g) (:127)
  solvePoissonEqn(psi, rho, ux, uy, uz)    solvePoissonEqn (:347-370): source (+ beta), solve
  makeVelocityDivFree(psi, rho, ux, ...)   momentumEqn.f90:795-869 (MG) / 999-1127 (FAST)
  updatePressure(psi, p, st)               updatePressure  (:374-459): p = psi, old levels rotated,
                                           clean_st_prev_bc (:419-459)
  setOldLevels(psi1, psi2, st1, st2)       the initial old levels of poissonEqnCTOR (:140-147) and of
                                           the VOF constructor (VOF.f90:186-191)

FAST_MODE keeps two old pressure levels, psi^n and psi^(n-1) (allocateOldField,
poissonEqn.f90:145-147), and two old surface-tension levels st^n, st^(n-1)
(allocateOldFieldV, VOF.f90:186), used by the splitting source and projection;
they start absent (zero), as in a run from rest, unless setOldLevels gives them.
Fields are CUDA tensors of the interior cells (shape (nz, ny, nx)); uy has ny+1
faces per column; a surface-tension field is a tuple (stx, sty, stz) laid out
like (ux, uy, uz).
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
        self.old_ = [None, None]  # psi^n, psi^(n-1)
        self.st_old_ = [None, None]  # st^n, st^(n-1): (stx, sty, stz) each
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

    def setOldLevels(self, psi1=None, psi2=None, st1=None, st2=None):
        """Copies of the initial old levels: psi%ptrf_ = psi and psi%ptrf_%ptrf_ = p0
        (poissonEqn.f90:140-147), st%ptrf_ = st and st%ptrf_%ptrf_ = st0
        (VOF.f90:186-191).  None leaves a level absent (zero)."""
        self.old_ = [None if a is None else a.clone() for a in (psi1, psi2)]
        self.st_old_ = [None if v is None else tuple(t.clone() for t in v) for v in (st1, st2)]

    def solvePoissonEqn(self, psi, rho, ux, uy, uz, alpha_dt, stream=None):
        """FAST: s = splitting source, solveFPS.  MG: s = div u*, beta = alpha dt / rho,
        solvePCG (psi is the warm start)."""
        s = self._work(rho)
        if self.mode == "FAST":
            F.computeSourceFast(self.grid, alpha_dt, self.rho0_, rho, ux, uy, uz, s, self.old_[0], self.old_[1],
                                self.st_old_[0], self.st_old_[1], stream=stream)
            self.fftSolver_.solveFPS(psi, s, stream=stream)
        else:
            F.computeSource(self.grid, ux, uy, uz, s, stream=stream)
            F.computeBeta(self.grid, alpha_dt, rho, self.beta_, stream=stream)
            self.pcgs_.solvePCG(psi, self.beta_, s, stream=stream)
        return psi

    def makeVelocityDivFree(self, psi, rho, ux, uy, uz, alpha_dt, stream=None):
        if self.mode == "FAST":
            F.makeVelocityDivFreeFast(self.grid, alpha_dt, self.rho0_, rho, psi, ux, uy, uz, self.old_[0],
                                      self.old_[1], self.st_old_[0], self.st_old_[1], stream=stream)
        else:
            F.makeVelocityDivFree(self.grid, alpha_dt, rho, psi, ux, uy, uz, stream=stream)

    def updatePressure(self, psi, p, st=None, stream=None):
        """p = psi; psi^(n-1) <- psi^n <- psi (storeOldField); st^(n-1) <- st^n <- st
        (storeOldFieldV) with the old st levels zeroed on the walls
        (clean_st_prev_bc).  Absent levels are created as zeros first; with st None
        and no st levels yet, the surface-tension levels stay absent."""
        import torch

        for a in range(2):
            if self.old_[a] is None:
                self.old_[a] = torch.zeros_like(psi)
        if st is not None:
            for a in range(2):
                if self.st_old_[a] is None:
                    self.st_old_[a] = tuple(torch.zeros_like(t) for t in st)
        with_st = self.st_old_[0] is not None and self.st_old_[1] is not None
        F.updatePressure(self.grid, psi, p, self.old_[0], self.old_[1], st if with_st else None,
                         self.st_old_[0] if with_st else None, self.st_old_[1] if with_st else None, stream=stream)

    def close(self):
        if self.mode == "FAST":
            self.fftSolver_.close()
        else:
            self.pcgs_.close()
