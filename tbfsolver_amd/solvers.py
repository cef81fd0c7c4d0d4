"""Host-side mirror of the reference's two pressure-solver plug-ins.

  FastPoissonSolver  <->  type(fastPoissonSolver), fastPoissonSolverCTOR, solveFPS
                          (src/fastPoissonSolver/fastPoissonSolver.f90:34-37,123-156,337-384)
  PCG                <->  type(pcg), pcgCTOR, solvePCG, iter_, res_
                          (src/pcg/pcg.f90:26-43,62-99,103-140)

Both accept either device tensors (torch CUDA tensors holding the interior
cells, Fortran order, i.e. a contiguous tensor of shape (nz, ny, nx)) or host
numpy arrays in Fortran order with optional halo layers (lower bounds given by
`lb`, default 1 - halo).  Every solve runs on the GPU through libtbfgpu.so;
there is no other code path.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import GridDesc, FpsOpts, PcgParams, check, lib, ptr, stream_ptr


def channel_grid(n, L=(2 * np.pi, 2.0, np.pi), stretch_y=False, gamma=1.4):
    """Channel grid: x, z periodic and uniform; y bounded by walls, uniform or
    tanh-stretched (grid.f90:293-330; the reference fixes gamma = 1.4)."""
    g = GridDesc()
    for a in range(3):
        g.n[a] = int(n[a])
        g.L[a] = float(L[a])
    g.stretch[0], g.stretch[1], g.stretch[2] = 0, int(bool(stretch_y)), 0
    g.gamma = float(gamma)
    return g


def _bounds(arr, lb, n):
    shape = arr.shape
    if lb is None:
        lb = tuple(1 - (shape[a] - n[a]) // 2 for a in range(3))
    lo = (C.c_int64 * 3)(*[int(v) for v in lb])
    hi = (C.c_int64 * 3)(*[int(lb[a]) + shape[a] - 1 for a in range(3)])
    return lo, hi


def _host(arr, writable=False):
    if not isinstance(arr, np.ndarray) or arr.dtype != np.float64 or not arr.flags.f_contiguous:
        raise TypeError("host arrays must be float64 numpy arrays in Fortran order")
    if writable and not arr.flags.writeable:
        raise TypeError("output array is read-only")
    return arr.ctypes.data_as(C.POINTER(C.c_double))


def _is_device(x):
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _check_dev(x, n, dtype_name="float64"):
    if not _is_device(x):
        raise TypeError("expected a CUDA tensor")
    if str(x.dtype) != "torch." + dtype_name or not x.is_contiguous() or x.numel() != n:
        raise TypeError("device tensors must be contiguous %s with %d elements" % (dtype_name, n))


class FastPoissonSolver:
    """FAST_MODE direct solver: FFT in x and z, tridiagonal in y."""

    def __init__(self, grid, precision="fp64", refine_sweeps=2):
        # fastPoissonSolverCTOR
        self.grid = grid
        o = FpsOpts()
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        o.precision = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
        o.refine_sweeps = int(refine_sweeps)
        h = C.c_void_p()
        check(lib().tbfg_fps_create(C.byref(grid), C.byref(o), C.byref(h)))
        self._h = h

    def solveFPS(self, p, s, lb_p=None, lb_s=None, stream=None):
        """Solves lap(p) = s; p is made volume-weighted mean free and, for host
        arrays, its halo is filled (fastPoissonSolver.f90:337-384)."""
        n = self.grid.shape
        if _is_device(p):
            _check_dev(p, self.grid.cells)
            _check_dev(s, self.grid.cells)
            check(lib().tbfg_fps_solve_dev(self._h, ptr(s), ptr(p), stream_ptr(stream)))
            return p
        slo, shi = _bounds(s, lb_s, n)
        plo, phi = _bounds(p, lb_p, n)
        check(lib().tbfg_fps_solve(self._h, _host(s), slo, shi, _host(p, True), plo, phi))
        return p

    def apply_operator(self, p, q, stream=None):
        """q = L p for the operator this solver inverts (device tensors)."""
        _check_dev(p, self.grid.cells)
        _check_dev(q, self.grid.cells)
        check(lib().tbfg_fps_apply_operator_dev(self._h, ptr(p), ptr(q), stream_ptr(stream)))
        return q

    @property
    def footprint(self):
        return int(lib().tbfg_fps_footprint(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().tbfg_fps_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_params():
    p = PcgParams()
    lib().tbfg_pcg_default_params(C.byref(p))
    return p


class PCG:
    """MG_MODE solver: PCG with a red/black Gauss-Seidel multigrid preconditioner.

    Parameters carry the reference's names (channel_18/specs/pcg_solver).  After
    solvePCG, `iter_` and `res_` hold what the reference's fields hold
    (pcg.f90:36-37: iter_ counts stop tests, i.e. updates + 1 on convergence) and
    `iterations` the number of direction updates.

    precision: "fp64" (the reference's arithmetic), "mixed" (FP32 multigrid
    inside the FP64 iteration) or "mixed-bf16" (as "mixed", with the finest
    level's multigrid right-hand side and coefficients stored in bf16)."""

    def __init__(self, grid, nLevels=4, nPreSweep=2, nPostSweep=2, tolPCG=1e-8, maxIterPCG=100,
                 precision="fp64", mg_cycles=1, coarse_max_iter=-1, coarse_tol=0.0, flexible=None,
                 preconditioner="mg", tolMG=0.0, maxIterMG=100):
        # pcgCTOR + multiGridCTOR
        self.grid = grid
        p = default_params()
        p.n_levels, p.n_pre, p.n_post = int(nLevels), int(nPreSweep), int(nPostSweep)
        p.tol, p.max_iter, p.mg_cycles = float(tolPCG), int(maxIterPCG), int(mg_cycles)
        precs = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "mixed-bf16": _lib.PREC_MIXED_BF16}
        if precision not in precs:
            raise ValueError("precision must be one of %s" % ", ".join(sorted(precs)))
        p.precision = precs[precision]
        p.coarse_max_iter = int(coarse_max_iter)
        p.coarse_tol = float(coarse_tol)
        p.flexible = -1 if flexible is None else int(flexible)
        if preconditioner not in ("mg", "fft"):
            raise ValueError("preconditioner must be 'mg' or 'fft'")
        p.preconditioner = 1 if preconditioner == "fft" else 0
        # solveMG's loop (multiGrid.f90:165-182): V-cycles per application until
        # the finest residual is at most tolMG; 0 keeps mg_cycles fixed
        p.mg_tol, p.mg_max_iter = float(tolMG), int(maxIterMG)
        self.params = p
        h = C.c_void_p()
        check(lib().tbfg_pcg_create(C.byref(grid), C.byref(p), C.byref(h)))
        self._h = h
        self.iter_ = 0
        self.res_ = 0.0
        self.iterations = 0
        self.converged = False

    @classmethod
    def from_pcg_solver_file(cls, grid, path, **kw):
        """Constructs with the keys of a reference `pcg_solver` parameter file."""
        prm = read_parameter_file(path)
        args = dict(nLevels=int(prm.get("nLevels", 4)), nPreSweep=int(prm.get("nPreSweep", 2)),
                    nPostSweep=int(prm.get("nPostSweep", 2)), tolPCG=float(prm.get("tolPCG", 1e-8)),
                    maxIterPCG=int(prm.get("maxIterPCG", 100)), tolMG=float(prm.get("tolMG", 0.0)),
                    maxIterMG=int(prm.get("maxIterMG", 100)))
        args.update(kw)
        return cls(grid, **args)

    def solvePCG(self, psi, beta, q, lb_psi=None, lb_beta=None, lb_q=None, stream=None):
        """Solves div(beta grad psi) = q from the warm start psi (pcg.f90:103-140)."""
        n = self.grid.shape
        it = C.c_int32(0)
        res = C.c_double(0.0)
        if _is_device(psi):
            for x in (psi, beta, q):
                _check_dev(x, self.grid.cells)
            rc = lib().tbfg_pcg_solve_dev(self._h, ptr(psi), ptr(beta), ptr(q), C.byref(it), C.byref(res),
                                          stream_ptr(stream))
        else:
            a, b = _bounds(psi, lb_psi, n)
            c, d = _bounds(beta, lb_beta, n)
            e, f = _bounds(q, lb_q, n)
            rc = lib().tbfg_pcg_solve(self._h, _host(psi, True), a, b, _host(beta), c, d, _host(q), e, f,
                                      C.byref(it), C.byref(res))
        check(rc, allow=(_lib.TBFG_ENOTCONVERGED,))
        self.iterations = int(it.value)
        self.res_ = float(res.value)
        self.converged = rc == _lib.TBFG_OK
        self.iter_ = self.iterations + 1 if self.converged else self.iterations
        return psi

    def solvePCG_queued(self, psi, beta, q, iterations, stream=None):
        """solvePCG with `iterations` iterations enqueued and no host synchronisation
        (device tensors; capturable in a graph).  The device's stop test still ends the
        solve; result() reads the outcome once the stream's work has finished."""
        for x in (psi, beta, q):
            _check_dev(x, self.grid.cells)
        check(lib().tbfg_pcg_solve_queued_dev(self._h, ptr(psi), ptr(beta), ptr(q), int(iterations),
                                              stream_ptr(stream)))
        return psi

    def result(self):
        """The outcome of the last queued solve: sets iterations, res_, converged, iter_."""
        it = C.c_int32(0)
        res = C.c_double(0.0)
        rc = lib().tbfg_pcg_result(self._h, C.byref(it), C.byref(res))
        check(rc, allow=(_lib.TBFG_ENOTCONVERGED,))
        self.iterations = int(it.value)
        self.res_ = float(res.value)
        self.converged = rc == _lib.TBFG_OK
        self.iter_ = self.iterations + 1 if self.converged else self.iterations
        return self.converged

    def precondition(self, beta, r, z, stream=None):
        """z = M r, one application of the multigrid preconditioner (device tensors)."""
        for x in (beta, r, z):
            _check_dev(x, self.grid.cells)
        check(lib().tbfg_pcg_precondition_dev(self._h, ptr(beta), ptr(r), ptr(z), stream_ptr(stream)))
        return z

    @property
    def footprint(self):
        return int(lib().tbfg_pcg_footprint(self._h))

    KERNELS = {"sweep_fine": 0, "restrict_fine": 1, "spmv": 2, "update": 3}

    def coarse_stats(self):
        """Coarsest-level solves of the last solve: dict(calls, iters, iters_max, res_max)."""
        a, b, c, d = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        check(lib().tbfg_pcg_coarse_stats(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"calls": a.value, "iters": b.value, "iters_max": c.value, "res_max": d.value}

    def mg_stats(self):
        """V-cycles of the last solve's preconditioner applications: (total, most in one)."""
        a, b = C.c_int32(0), C.c_int32(0)
        check(lib().tbfg_pcg_mg_stats(self._h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def profile_begin(self, kernel):
        check(lib().tbfg_pcg_profile_begin(self._h, self.KERNELS.get(kernel, kernel)))

    def profile_end(self):
        """(launches, total_ms, algorithmic bytes per launch) since profile_begin."""
        n, ms, b = C.c_int64(0), C.c_double(0.0), C.c_double(0.0)
        check(lib().tbfg_pcg_profile_end(self._h, C.byref(n), C.byref(ms), C.byref(b)))
        return int(n.value), float(ms.value), float(b.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().tbfg_pcg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_parameter_file(path):
    """Reads a reference `specs/` parameter file: `name value` per line, `!`
    comments (parFile.f90:60-259).  Fortran literals like 1.d-10 and .T. are
    converted."""
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.split("!", 1)[0].strip()
            if not line:
                continue
            parts = line.split()
            if len(parts) < 2:
                continue
            key, val = parts[0], parts[1]
            v = val.lower()
            if v in (".t.", ".true."):
                out[key] = True
            elif v in (".f.", ".false."):
                out[key] = False
            else:
                try:
                    out[key] = int(val)
                except ValueError:
                    try:
                        out[key] = float(v.replace("d", "e"))
                    except ValueError:
                        out[key] = val
    return out
