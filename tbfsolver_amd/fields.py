"""Device-resident pieces around the pressure solve, with the reference's names:

  computeSource (MG)      poissonEqn.f90:155-190      div(u*) of the staggered velocity
  computeBeta             poissonEqn.f90:311-342      beta = alpha dt / rho
  makeVelocityDivFree     momentumEqn.f90:795-869,980-993
  computeSourceFast       poissonEqn.f90:193-305      FAST_MODE source (rho0 splitting, old pressure)
  makeVelocityDivFreeFast momentumEqn.f90:999-1127
  updatePressure          poissonEqn.f90:374-459      p = psi, old pressure and surface-tension levels
  computeContinuityError  auxiliaryRoutines.f90:42-91 max|div u| * dt

and the host-side input mirrors used by the tests and benchmarks (seeded
staggered velocity; bubble density from spheres, initialConditions.f90:192-398
and VOF.f90:197-262).

Staggered arrays: ux, uz have nx*ny*nz faces (tensor shape (nz, ny, nx)); uy
has nx*(ny+1)*nz faces (shape (nz, ny+1, nx)) with the walls at j = 0 and ny.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib, ptr, ptr3, stream_ptr


def computeSource(grid, ux, uy, uz, s, stream=None):
    check(lib().tbfg_divergence_dev(C.byref(grid), ptr(ux), ptr(uy), ptr(uz), ptr(s), stream_ptr(stream)))
    return s


def computeBeta(grid, alpha_dt, rho, beta, stream=None):
    check(lib().tbfg_beta_dev(C.byref(grid), float(alpha_dt), ptr(rho), ptr(beta), stream_ptr(stream)))
    return beta


def makeVelocityDivFree(grid, alpha_dt, rho, psi, ux, uy, uz, stream=None):
    check(lib().tbfg_project_mg_dev(C.byref(grid), float(alpha_dt), ptr(rho), ptr(psi), ptr(ux), ptr(uy),
                                    ptr(uz), stream_ptr(stream)))


def computeSourceFast(grid, alpha_dt, rho0, rho, ux, uy, uz, s, psi1=None, psi2=None, st1=None, st2=None,
                      stream=None):
    """FAST_MODE source with the old-pressure splitting (poissonEqn.f90:193-305);
    psi1 = psi^n, psi2 = psi^(n-1); st1 = st^n, st2 = st^(n-1), each a tuple
    (stx, sty, stz) of staggered tensors laid out like (ux, uy, uz).  None before
    they exist (zero)."""
    check(lib().tbfg_source_fast_dev(C.byref(grid), float(alpha_dt), float(rho0), ptr(rho), ptr(ux), ptr(uy),
                                     ptr(uz), ptr(psi1), ptr(psi2), ptr3(st1), ptr3(st2), ptr(s),
                                     stream_ptr(stream)))
    return s


def makeVelocityDivFreeFast(grid, alpha_dt, rho0, rho, psi, ux, uy, uz, psi1=None, psi2=None, st1=None, st2=None,
                            stream=None):
    """FAST_MODE projection (momentumEqn.f90:999-1127)."""
    check(lib().tbfg_project_fast_dev(C.byref(grid), float(alpha_dt), float(rho0), ptr(rho), ptr(psi), ptr(psi1),
                                      ptr(psi2), ptr3(st1), ptr3(st2), ptr(ux), ptr(uy), ptr(uz),
                                      stream_ptr(stream)))


def updatePressure(grid, psi, p, psi1, psi2, st=None, st1=None, st2=None, stream=None):
    """updatePressure + clean_st_prev_bc (poissonEqn.f90:374-459) in one kernel:
    p = psi, psi2 <- psi1 <- psi, st2 <- st1 <- st, old st levels zero on the
    walls.  p may be None; st None with st1, st2 given stores zero surface tension."""
    check(lib().tbfg_update_pressure_dev(C.byref(grid), ptr(psi), ptr(p), ptr(psi1), ptr(psi2), ptr3(st),
                                         ptr3(st1), ptr3(st2), stream_ptr(stream)))


def updateMaterialProps(grid, c, cs, rho, mu=None, rho_l=1.0, rho_g=1.0 / 20.0, mu_l=1.0, mu_g=1.0 / 50.0,
                        work=None, stream=None):
    """updateMaterialProps (VOF.f90:197-229): cs = smoothVF(c), rho and mu from
    cs.  work: a device tensor of tbfg_material_work_doubles doubles (allocated
    here when None)."""
    if work is None:
        import torch

        work = torch.empty(int(lib().tbfg_material_work_doubles(C.byref(grid))), dtype=torch.float64,
                           device=c.device)
    check(lib().tbfg_material_props_dev(C.byref(grid), float(rho_l), float(rho_g), float(mu_l), float(mu_g), ptr(c),
                                        ptr(cs), ptr(rho), ptr(mu) if mu is not None else None, ptr(work),
                                        stream_ptr(stream)))
    return cs


def max_divergence(grid, ux, uy, uz, stream=None):
    out = C.c_double(0.0)
    check(lib().tbfg_max_divergence_dev(C.byref(grid), ptr(ux), ptr(uy), ptr(uz), C.byref(out),
                                        stream_ptr(stream)))
    return float(out.value)


def max_divergence_noalloc(grid, ux, uy, uz, stream=None):
    """max_divergence without any device allocation (static scratch, axes <= 4096 cells)."""
    out = C.c_double(0.0)
    check(lib().tbfg_max_divergence_noalloc_dev(C.byref(grid), ptr(ux), ptr(uy), ptr(uz), C.byref(out),
                                                stream_ptr(stream)))
    return float(out.value)


def computeContinuityError(grid, ux, uy, uz, dt, stream=None):
    """max|div u| * dt, as the reference prints it."""
    return max_divergence(grid, ux, uy, uz, stream) * dt


def apply_operator(grid, beta, p, q, stream=None):
    """q = div(beta grad p) with the channel ghosts (computeQv, pcg.f90:144-183)."""
    check(lib().tbfg_apply_operator_dev(C.byref(grid), ptr(beta), ptr(p), ptr(q), stream_ptr(stream)))
    return q


def make_velocity(grid, seed):
    """Host arrays (Fortran order): ux, uz of shape (nx, ny, nz), uy of shape (nx, ny+1, nz)."""
    nx, ny, nz = grid.shape
    ux = np.zeros((nx, ny, nz), order="F")
    uz = np.zeros((nx, ny, nz), order="F")
    uy = np.zeros((nx, ny + 1, nz), order="F")
    d = C.POINTER(C.c_double)
    check(lib().tbfg_make_velocity(C.byref(grid), C.c_uint64(seed), ux.ctypes.data_as(d),
                                   uy.ctypes.data_as(d), uz.ctypes.data_as(d)))
    return ux, uy, uz


def make_bubble_density(grid, nb, diameter, seed, nref=50, rho_l=1.0, rho_g=1.0 / 20.0):
    """Host density field (Fortran order) and the bubble centres (nb x 3)."""
    rho = np.zeros(grid.shape, order="F")
    cen = np.zeros((max(nb, 1), 3))
    d = C.POINTER(C.c_double)
    check(lib().tbfg_make_bubble_density(C.byref(grid), int(nb), float(diameter), C.c_uint64(seed), int(nref),
                                         float(rho_l), float(rho_g), rho.ctypes.data_as(d), cen.ctypes.data_as(d)))
    return rho, cen[:nb]
