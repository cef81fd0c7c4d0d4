"""ctypes front-end of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
CPU-baseline leg as the checker.  The product never imports this module.

Arrays are numpy float64, Fortran-ordered (shape (nx, ny, nz), i fastest), the
layout the GPU library's C ABI uses for device-resident fields.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class GridDesc(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("L", C.c_double * 3),
                ("stretch", C.c_int32 * 3), ("gamma", C.c_double)]


class PcgParams(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("n_pre", C.c_int32), ("n_post", C.c_int32),
                ("max_iter", C.c_int32), ("tol", C.c_double), ("mg_cycles", C.c_int32),
                ("flexible", C.c_int32), ("coarse_mode", C.c_int32),
                ("coarse_cycles", C.c_int32), ("mg_max_iter", C.c_int32), ("mg_tol", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        _LIB = C.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f(a):
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray(a)


def desc(n, L=(2 * np.pi, 2.0, np.pi), stretch=(0, 0, 0), gamma=1.4):
    g = GridDesc()
    for a in range(3):
        g.n[a] = int(n[a])
        g.L[a] = float(L[a])
        g.stretch[a] = int(stretch[a])
    g.gamma = float(gamma)
    return g


def params(n_levels=4, n_pre=2, n_post=2, max_iter=200, tol=1e-8, mg_cycles=1, flexible=0,
           coarse_mode=0, coarse_cycles=4, mg_tol=0.0, mg_max_iter=100):
    p = PcgParams()
    p.n_levels, p.n_pre, p.n_post, p.max_iter = n_levels, n_pre, n_post, max_iter
    p.tol, p.mg_cycles, p.flexible = tol, mg_cycles, flexible
    p.coarse_mode, p.coarse_cycles = coarse_mode, coarse_cycles
    p.mg_tol, p.mg_max_iter = mg_tol, mg_max_iter
    return p


def last_mg_cycles(with_last=False):
    """(total, largest per application[, last application]) V-cycles of the
    last pcg_solve / mg_apply."""
    a, b, c = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    lib().or_last_mg_cycles(C.byref(a), C.byref(b), C.byref(c))
    return (int(a.value), int(b.value), int(c.value)) if with_last else (int(a.value), int(b.value))


def shape(g, lev=0):
    return tuple(int(g.n[a]) >> lev for a in range(3))


def axis_coords(n, L, stretch, gamma, ghost):
    xc = np.zeros(n + 2 * ghost)
    xf = np.zeros(n + 2 * ghost + 1)
    lib().or_axis_coords(C.c_int32(n), C.c_double(L), C.c_int32(stretch), C.c_double(gamma),
                         C.c_int32(ghost), _p(xc), _p(xf))
    return xc, xf


def fft_c2c(x, sign):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    out = np.zeros_like(x)
    lib().or_fft_c2c(C.c_int32(x.size), x.ctypes.data_as(C.POINTER(C.c_double)),
                     out.ctypes.data_as(C.POINTER(C.c_double)), C.c_int32(sign))
    return out


def apply_op(g, beta, p):
    beta, p = _f(beta), _f(p)
    q = np.zeros_like(p, order="F")
    lib().or_apply_op(C.byref(g), _p(beta), _p(p), _p(q))
    return q


def fft_solve(g, s):
    s = _f(s)
    p = np.zeros_like(s, order="F")
    rc = lib().or_fft_solve(C.byref(g), _p(s), _p(p))
    if rc != 0:
        raise ValueError("or_fft_solve rejected the grid")
    return p


def fft_operator(g, p):
    p = _f(p)
    q = np.zeros_like(p, order="F")
    lib().or_fft_operator(C.byref(g), _p(p), _p(q))
    return q


def tridiag(g, lam, pinned, rhs):
    d = np.ascontiguousarray(np.asarray(rhs, dtype=np.complex128)).copy()
    lib().or_tridiag(C.byref(g), C.c_double(lam), C.c_int32(pinned),
                     d.ctypes.data_as(C.POINTER(C.c_double)))
    return d


def rbgs(g, lev, beta, q, p, sweeps, remove_mean=1):
    beta, q = _f(beta), _f(q)
    p = _f(p).copy(order="F")
    lib().or_rbgs(C.byref(g), C.c_int32(lev), _p(beta), _p(q), _p(p), C.c_int32(sweeps),
                  C.c_int32(remove_mean))
    return p


def residual(g, lev, beta, q, p):
    beta, q, p = _f(beta), _f(q), _f(p)
    r = np.zeros_like(p, order="F")
    lib().or_residual(C.byref(g), C.c_int32(lev), _p(beta), _p(q), _p(p), _p(r))
    return r


def restrict(g, lev_fine, fine):
    fine = _f(fine)
    out = np.zeros(shape(g, lev_fine + 1), order="F")
    lib().or_restrict(C.byref(g), C.c_int32(lev_fine), _p(fine), _p(out))
    return out


def prolong_add(g, lev_fine, coarse, fine):
    coarse = _f(coarse)
    fine = _f(fine).copy(order="F")
    lib().or_prolong_add(C.byref(g), C.c_int32(lev_fine), _p(coarse), _p(fine))
    return fine


def coarse_exact(g, lev, beta, q):
    beta, q = _f(beta), _f(q)
    p = np.zeros_like(q, order="F")
    lib().or_coarse_exact(C.byref(g), C.c_int32(lev), _p(beta), _p(q), _p(p))
    return p


def mg_apply(g, prm, beta, r):
    beta, r = _f(beta), _f(r)
    z = np.zeros_like(r, order="F")
    rc = lib().or_mg_apply(C.byref(g), C.byref(prm), _p(beta), _p(r), _p(z))
    if rc != 0:
        raise ValueError("or_mg_apply rejected the hierarchy")
    return z


def pcg_solve(g, prm, beta, q, psi0=None):
    beta, q = _f(beta), _f(q)
    psi = np.zeros_like(q, order="F") if psi0 is None else _f(psi0).copy(order="F")
    it = C.c_int32(0)
    res = C.c_double(0.0)
    rc = lib().or_pcg_solve(C.byref(g), C.byref(prm), _p(beta), _p(q), _p(psi), C.byref(it),
                            C.byref(res))
    if rc < 0:
        raise ValueError("or_pcg_solve rejected the hierarchy")
    return psi, int(it.value), float(res.value), int(rc)


def divergence(g, ux, uy, uz):
    ux, uy, uz = _f(ux), _f(uy), _f(uz)
    d = np.zeros(shape(g), order="F")
    lib().or_divergence(C.byref(g), _p(ux), _p(uy), _p(uz), _p(d))
    return d


def beta_from_rho(g, alpha_dt, rho):
    rho = _f(rho)
    b = np.zeros_like(rho, order="F")
    lib().or_beta(C.byref(g), C.c_double(alpha_dt), _p(rho), _p(b))
    return b


def project_mg(g, alpha_dt, rho, psi, ux, uy, uz):
    rho, psi = _f(rho), _f(psi)
    ux, uy, uz = (_f(a).copy(order="F") for a in (ux, uy, uz))
    lib().or_project_mg(C.byref(g), C.c_double(alpha_dt), _p(rho), _p(psi), _p(ux), _p(uy), _p(uz))
    return ux, uy, uz


def _levels(st):
    """A surface-tension level (ux-, uy-, uz-shaped arrays) as a C array of three
    pointers and the arrays it points into; None is a NULL level (zero)."""
    if st is None:
        return None, ()
    arrs = tuple(_f(a) for a in st)
    return (C.POINTER(C.c_double) * 3)(*(_p(a) for a in arrs)), arrs


def project_fast(g, alpha_dt, rho0, rho, psi, psi1, psi2, ux, uy, uz, st1=None, st2=None):
    rho, psi, psi1, psi2 = _f(rho), _f(psi), _f(psi1), _f(psi2)
    ux, uy, uz = (_f(a).copy(order="F") for a in (ux, uy, uz))
    l1, keep1 = _levels(st1)
    l2, keep2 = _levels(st2)
    lib().or_project_fast(C.byref(g), C.c_double(alpha_dt), C.c_double(rho0), _p(rho), _p(psi),
                          _p(psi1), _p(psi2), l1, l2, _p(ux), _p(uy), _p(uz))
    return ux, uy, uz


def source_fast(g, alpha_dt, rho0, rho, ux, uy, uz, psi1, psi2, st1=None, st2=None):
    rho, ux, uy, uz, psi1, psi2 = (_f(a) for a in (rho, ux, uy, uz, psi1, psi2))
    s = np.zeros(shape(g), order="F")
    l1, keep1 = _levels(st1)
    l2, keep2 = _levels(st2)
    lib().or_source_fast(C.byref(g), C.c_double(alpha_dt), C.c_double(rho0), _p(rho), _p(ux),
                         _p(uy), _p(uz), _p(psi1), _p(psi2), l1, l2, _p(s))
    return s


def update_pressure(g, psi, psi1, psi2, st=None, st1=None, st2=None):
    """updatePressure with clean_st_prev_bc: returns (p, psi1, psi2, st1, st2),
    new arrays (st1, st2 None when they were)."""
    psi = _f(psi)
    p = np.zeros_like(psi, order="F")
    psi1, psi2 = _f(psi1).copy(order="F"), _f(psi2).copy(order="F")
    if st1 is not None:
        st1 = tuple(_f(a).copy(order="F") for a in st1)
        st2 = tuple(_f(a).copy(order="F") for a in st2)
    lc, keepc = _levels(st)
    l1, _ = _levels(st1)
    l2, _ = _levels(st2)
    lib().or_update_pressure(C.byref(g), _p(psi), _p(p), _p(psi1), _p(psi2), lc, l1, l2)
    return p, psi1, psi2, st1, st2


def material_props(g, c, rho_l=1.0, rho_g=1.0 / 20.0, mu_l=1.0, mu_g=1.0 / 50.0):
    """(cs, rho, mu): the smoothed volume fraction and the material properties."""
    c = _f(c)
    cs, rho, mu = (np.zeros(shape(g), order="F") for _ in range(3))
    lib().or_material_props(C.byref(g), C.c_double(rho_l), C.c_double(rho_g), C.c_double(mu_l), C.c_double(mu_g),
                            _p(c), _p(cs), _p(rho), _p(mu))
    return cs, rho, mu


def volume_mean(g, p):
    p = _f(p)
    lib().or_volume_mean.restype = C.c_double
    return float(lib().or_volume_mean(C.byref(g), _p(p)))
