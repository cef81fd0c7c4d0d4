"""GPU parity of the FAST_MODE path (FFT x/z + tridiagonal y) against the CPU oracle.

FP64: the pressure agrees with the oracle to a relative L2 error of 1e-12 and
the projected velocity of config 0 has max|div u| <= 1e-10.  Mixed precision
(FP32 transforms, FP64 (0,0) column and FP64 refinement): relative L2 error
against the FP64 oracle <= 1e-10 after two refinement sweeps.
"""
import numpy as np
import pytest

import tbfsolver_amd as T
from tbfsolver_amd import fields as F
from oracle import pyoracle as O
from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(99)


def grids(n, stretch=0, gamma=1.5):
    return T.channel_grid(n, stretch_y=stretch, gamma=gamma), O.desc(n, stretch=(0, stretch, 0), gamma=gamma)


def vol(go):
    n = O.shape(go)
    f = [np.diff(O.axis_coords(n[a], go.L[a], go.stretch[a], go.gamma, 0)[1]) for a in range(3)]
    return f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# (lines of 32, 64, 96, 128, 192 or 256 points run the two-level passes where
# the other axis tiles into their workgroups, the others the Stockham passes:
# (32, 24, 16), (96, 32, 48) and (192, 128, 96) mix the two.  (1024, 8, 768):
# the longest lines of the default LDS tiling; (16, 1024, 8): the tallest y
# column, 16 rows per lane of the tridiagonal solve; the last three: lines of
# 2048, 3072 and 4096 points, one line per workgroup with more than 64 KB of LDS;
# (32, 256, 16) and (16, 512, 8): 4 and 8 rows per lane of the y solve's form for
# ny = 64 x rows per lane; (32, 96, 16) and (32, 160, 16): 2 and 3 rows per lane
# with padded rows in the last lanes)
CASES = [((64, 64, 64), 0), ((32, 24, 16), 1), ((96, 32, 48), 1), ((256, 192, 256), 1), ((192, 128, 96), 0),
         ((256, 16, 32), 1), ((32, 16, 256), 0), ((8, 6, 4), 1), ((1024, 8, 768), 0), ((16, 1024, 8), 1),
         ((2048, 8, 4), 1), ((3072, 4, 2), 0), ((6, 4, 4096), 1), ((32, 256, 16), 1), ((16, 512, 8), 0),
         ((32, 96, 16), 1), ((32, 160, 16), 1), ((384, 4, 144), 0), ((72, 36, 96), 1),
         ((96, 4, 576), 1), ((12, 12, 36), 1)]


@pytest.mark.parametrize("n,stretch", CASES)
def test_fft_solve_fp64(torch_gpu, n, stretch):
    print("case fft fp64", n, stretch, flush=True)
    g, go = grids(n, stretch)
    V = vol(go)
    s = rng.standard_normal(n)
    s -= np.sum(V * s) / V.sum()
    fs = T.FastPoissonSolver(g)
    p = torch_gpu.zeros_like(to_dev(s, torch_gpu))
    fs.solveFPS(p, to_dev(s, torch_gpu))
    torch_gpu.cuda.synchronize()
    out = to_host(p)
    # round-off grows with the conditioning of the y solve, ~ ny^2 (1024 stretched
    # rows: 1.7e-12 forward, 3.5e-12 backward against 1e-12 up to 256 rows)
    scale = max(1.0, (n[1] / 256) ** 2)
    if np.prod(n) <= 300000:
        ref = O.fft_solve(go, s)
        assert rel(out, ref) < 1e-12 * scale
    # residual of the operator the solver inverts, and the volume-weighted mean
    # (its FP64 round-off grows with the operator's norm, ~ (n/1024)^2 along x or z
    # lines longer than 1024 points: 2.3e-12 at 3072; and the y solve's partitioned
    # Thomas + PCR leaves 0.6-1.4e-12 at 256 stretched rows, where a sequential
    # Thomas sweep leaves 1e-13: a factor 2 from 256 rows)
    q = torch_gpu.zeros_like(p)
    fs.apply_operator(p, q)
    torch_gpu.cuda.synchronize()
    rscale = max(scale, (max(n[0], n[2]) / 1024) ** 2, 2.0 if n[1] >= 256 else 1.0)
    assert np.linalg.norm(to_host(q) - s) < 1e-12 * rscale * np.linalg.norm(s)
    assert abs(np.sum(V * out)) / V.sum() < 1e-13 * np.abs(out).max()


@pytest.mark.parametrize("n,stretch", [((64, 64, 64), 0), ((96, 32, 48), 1), ((256, 16, 32), 1), ((256, 192, 256), 1),
                                       ((2048, 8, 4), 1), ((72, 36, 96), 1), ((384, 4, 576), 0)])
def test_fft_solve_mixed(torch_gpu, n, stretch):
    print("case fft mixed", n, stretch, flush=True)
    g, go = grids(n, stretch)
    V = vol(go)
    s = rng.standard_normal(n)
    s -= np.sum(V * s) / V.sum()
    ref_solver = T.FastPoissonSolver(g)
    pr = torch_gpu.zeros_like(to_dev(s, torch_gpu))
    ref_solver.solveFPS(pr, to_dev(s, torch_gpu))
    for sweeps, bound in ((0, 1e-3), (1, 1e-7), (2, 1e-10)):
        fs = T.FastPoissonSolver(g, precision="mixed", refine_sweeps=sweeps)
        p = torch_gpu.zeros_like(pr)
        fs.solveFPS(p, to_dev(s, torch_gpu))
        torch_gpu.cuda.synchronize()
        e = rel(to_host(p), to_host(pr))
        print("  sweeps", sweeps, "rel err", e, flush=True)
        assert e < bound


def test_fft_lengths_beyond_the_passes_are_rejected(torch_gpu):
    """Lines the FFT passes cannot take give EINVAL at creation: 3456 = 4^3 2 3^3
    points (its radix-2 pass has 1728 butterflies, more than a workgroup's 1024 in
    FP64) and 4098 points (beyond 4096)."""
    from tbfsolver_amd import _lib
    for n in ((3456, 4, 2), (4098, 4, 2)):
        with pytest.raises(_lib.TbfgError) as e:
            T.FastPoissonSolver(T.channel_grid(n), precision="fp64")
        assert e.value.code == _lib.TBFG_EINVAL, n


def test_fft_solve_full_size_against_oracle(torch_gpu):
    """The bench grid (256x192x256, tanh y, two-level passes along x and z): the
    pressure against the oracle's FP64 solve of the same right-hand side, FP64
    within 1e-11 and FP32 transforms with two FP64 refinement sweeps within 1e-10
    (relative L2; the FP64 difference, 3.9e-12, is round-off of the two solves
    through the stretched 192-row y columns, the operator residual of the same
    solve staying below 1e-12)."""
    n = (256, 192, 256)
    print("case fft full size vs oracle", n, flush=True)
    g, go = grids(n, 1)
    V = vol(go)
    s = np.random.default_rng(2024).standard_normal(n)
    s -= np.sum(V * s) / V.sum()
    ref = O.fft_solve(go, s)
    for precision, bound in (("fp64", 1e-11), ("mixed", 1e-10)):
        fs = T.FastPoissonSolver(g, precision=precision)
        p = torch_gpu.zeros(n[::-1], dtype=torch_gpu.float64, device="cuda")
        fs.solveFPS(p, to_dev(s, torch_gpu))
        torch_gpu.cuda.synchronize()
        e = rel(to_host(p), ref)
        print("  %s rel err %.2e" % (precision, e), flush=True)
        assert e < bound
        fs.close()


@pytest.mark.parametrize("n", [(256, 48, 256), (128, 32, 192), (96, 16, 64), (32, 8, 32)])
@pytest.mark.parametrize("precision,bound", [("fp64", 1e-13), ("mixed", 1e-5)])
def test_two_level_passes_equal_stockham_passes(torch_gpu, monkeypatch, n, precision, bound):
    """Every two-level split (16x16, 16x8 and 16x12, 12x8 and 8x8, 8x4) against
    the Stockham passes of the same lengths (TBFG_FFT_TWO_LEVEL=0, read when a
    handle is created), on the same right-hand side; mixed precision without
    refinement compares the FP32 transforms themselves."""
    print("case two-level vs stockham", precision, n, flush=True)
    g, go = grids(n, 1)
    V = vol(go)
    s = rng.standard_normal(n)
    s -= np.sum(V * s) / V.sum()
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("TBFG_FFT_TWO_LEVEL", flag)
        fs = T.FastPoissonSolver(g, precision=precision, refine_sweeps=0)
        p = torch_gpu.zeros(n[::-1], dtype=torch_gpu.float64, device="cuda")
        fs.solveFPS(p, to_dev(s, torch_gpu))
        torch_gpu.cuda.synchronize()
        out[flag] = to_host(p)
        fs.close()
    e = rel(out["1"], out["0"])
    print("  rel diff", e, flush=True)
    assert e < bound


def test_config0_projection(torch_gpu):
    """64^3 constant density: s = div(u)/dt, FFT solve, u -= dt grad p."""
    print("case config0 projection", flush=True)
    n = (64, 64, 64)
    g, go = grids(n, 0)
    dt = 1e-3
    ux, uy, uz = F.make_velocity(g, 1000)
    dux, duy, duz = (to_dev(a, torch_gpu) for a in (ux, uy, uz))
    s = torch_gpu.zeros_like(dux)
    F.computeSource(g, dux, duy, duz, s)
    s /= dt
    s -= s.mean()
    fs = T.FastPoissonSolver(g)
    p = torch_gpu.zeros_like(s)
    fs.solveFPS(p, s)
    ones = torch_gpu.ones_like(s)
    F.makeVelocityDivFree(g, dt, ones, p, dux, duy, duz)
    md = F.max_divergence(g, dux, duy, duz)
    assert md <= 1e-10, md
    # against the oracle's full config-0 step
    so = O.source_fast(go, dt, 1.0, np.ones(n), ux, uy, uz, np.zeros(n), np.zeros(n))
    so -= so.mean()
    po = O.fft_solve(go, so)
    assert rel(to_host(p), po) < 1e-10


def test_host_array_entry_point(torch_gpu):
    print("case fft host arrays", flush=True)
    n = (32, 24, 16)
    g, go = grids(n, 1)
    V = vol(go)
    s = rng.standard_normal(n)
    s -= np.sum(V * s) / V.sum()
    fs = T.FastPoissonSolver(g)
    p_h = np.zeros((34, 26, 18), order="F")
    fs.solveFPS(p_h, np.asfortranarray(np.pad(s, 1)))
    ref = O.fft_solve(go, s)
    assert rel(p_h[1:-1, 1:-1, 1:-1], ref) < 1e-12
    assert np.array_equal(p_h[-1, 1:-1, 1:-1], p_h[1, 1:-1, 1:-1])
    assert np.array_equal(p_h[1:-1, -1, 1:-1], p_h[1:-1, -2, 1:-1])


def max_div(go, dux, duy, duz, walls_clean=True):
    """max|div u| of a projected device velocity (host divergence).  The source reads the
    surface-tension levels on the y wall faces (computeOldPressDiv, poissonEqn.f90:273-291)
    but the projection never changes a wall face, so with non-zero st on the walls the two
    wall rows keep those terms, in the reference too; clean_st_prev_bc (:419-459) zeroes
    them for the following steps.  With walls_clean False only interior rows are taken."""
    d = O.divergence(go, to_host(dux), to_host(duy), to_host(duz))
    return np.abs(d if walls_clean else d[:, 1:-1, :]).max()


def random_st(n, walls=True):
    """A random staggered field laid out like (ux, uy, uz); with walls, the y wall
    faces are non-zero too (the reference reads them as stored)."""
    nx, ny, nz = n
    sx, sz = rng.standard_normal(n), rng.standard_normal(n)
    sy = rng.standard_normal((nx, ny + 1, nz))
    if not walls:
        sy[:, 0, :] = 0
        sy[:, -1, :] = 0
    return sx, sy, sz


@pytest.mark.parametrize("levels", ["none", "psi", "psi+st", "psi+st+walls"])
def test_fast_source_and_projection_variable_density(torch_gpu, levels):
    """FAST_MODE source and projection with the rho0 splitting (poissonEqn.f90:193-305,
    momentumEqn.f90:999-1127) on a stretched channel with bubbles, against the oracle;
    then the full step s -> solveFPS -> projection, compared field by field.  With
    "psi+st" the two old surface-tension levels are non-zero on every face but the y
    walls (as updatePressure leaves them); with "psi+st+walls" on the wall faces too."""
    n = (64, 48, 32)
    print("case fast source/projection", n, levels, flush=True)
    g, go = grids(n, 1)
    rho, _ = F.make_bubble_density(g, 6, 0.3, 1003, nref=8)
    ux, uy, uz = F.make_velocity(g, 1003)
    rho0, adt = 1.0 / 20.0, 2e-3
    old = levels != "none"
    psi1 = rng.standard_normal(n) if old else np.zeros(n)
    psi2 = rng.standard_normal(n) if old else np.zeros(n)
    walls = levels == "psi+st+walls"
    st1 = random_st(n, walls) if levels.startswith("psi+st") else None
    st2 = random_st(n, walls) if levels.startswith("psi+st") else None
    d = {k: to_dev(v, torch_gpu) for k, v in dict(rho=rho, ux=ux, uy=uy, uz=uz, psi1=psi1, psi2=psi2).items()}
    dst1 = None if st1 is None else tuple(to_dev(a, torch_gpu) for a in st1)
    dst2 = None if st2 is None else tuple(to_dev(a, torch_gpu) for a in st2)
    s = torch_gpu.zeros_like(d["rho"])
    F.computeSourceFast(g, adt, rho0, d["rho"], d["ux"], d["uy"], d["uz"], s,
                        d["psi1"] if old else None, d["psi2"] if old else None, dst1, dst2)
    so = O.source_fast(go, adt, rho0, rho, ux, uy, uz, psi1, psi2, st1, st2)
    assert rel(to_host(s), so) < 1e-13
    fs = T.FastPoissonSolver(g)
    p = torch_gpu.zeros_like(s)
    fs.solveFPS(p, s)
    po = O.fft_solve(go, so)
    assert rel(to_host(p), po) < 1e-12
    F.makeVelocityDivFreeFast(g, adt, rho0, d["rho"], p, d["ux"], d["uy"], d["uz"],
                              d["psi1"] if old else None, d["psi2"] if old else None, dst1, dst2)
    uxo, uyo, uzo = O.project_fast(go, adt, rho0, rho, po, psi1, psi2, ux, uy, uz, st1, st2)
    for dev_a, ref in ((d["ux"], uxo), (d["uy"], uyo), (d["uz"], uzo)):
        assert rel(to_host(dev_a), ref) < 1e-12
    assert np.all(to_host(d["uy"])[:, 0, :] == 0) and np.all(to_host(d["uy"])[:, -1, :] == 0)
    # the step leaves the velocity divergence free whatever the old levels are (on the
    # interior rows only when the st wall faces are non-zero)
    scale = np.abs(O.divergence(go, ux, uy, uz)).max()
    assert max_div(go, d["ux"], d["uy"], d["uz"], walls_clean=not walls) <= 1e-10 * scale
    if not walls:
        assert F.max_divergence(g, d["ux"], d["uy"], d["uz"]) <= 1e-10 * scale
    else:
        assert max_div(go, d["ux"], d["uy"], d["uz"]) > 1e-6 * scale   # the wall rows keep the st terms
    fs.close()


def test_fast_step_full_size(torch_gpu):
    """One FAST_MODE step at 256x192x256 (stretched y, 72 bubbles, both old
    pressure levels and both old surface-tension levels non-zero, walls clean):
    source, solve and projection each against the oracle (the solve and the
    velocities within the 1e-11 of the full-size solve), and the projected
    velocity divergence free to 1e-10 of the source's scale and to 1e-10 in
    absolute terms (raw max|div u|, not multiplied by dt; measured: pressure
    8.7e-14, velocities 1.5e-14, max|div u| 2.8e-11)."""
    n = (256, 192, 256)
    print("case fast step full size", n, flush=True)
    g, go = grids(n, 1)
    lr = np.random.default_rng(77)
    rho, _ = F.make_bubble_density(g, 72, 0.3, 1002, nref=10)
    ux, uy, uz = F.make_velocity(g, 1002)
    rho0, adt = 1.0 / 20.0, 1e-3
    psi1, psi2 = lr.standard_normal(n), lr.standard_normal(n)

    def st():
        sy = lr.standard_normal((n[0], n[1] + 1, n[2]))
        sy[:, 0, :] = 0
        sy[:, -1, :] = 0
        return lr.standard_normal(n), sy, lr.standard_normal(n)

    st1, st2 = st(), st()
    d = {k: to_dev(v, torch_gpu) for k, v in dict(rho=rho, ux=ux, uy=uy, uz=uz, psi1=psi1, psi2=psi2).items()}
    dst1, dst2 = (tuple(to_dev(a, torch_gpu) for a in x) for x in (st1, st2))
    s = torch_gpu.zeros_like(d["rho"])
    F.computeSourceFast(g, adt, rho0, d["rho"], d["ux"], d["uy"], d["uz"], s, d["psi1"], d["psi2"], dst1, dst2)
    so = O.source_fast(go, adt, rho0, rho, ux, uy, uz, psi1, psi2, st1, st2)
    assert rel(to_host(s), so) < 1e-13
    fs = T.FastPoissonSolver(g)
    p = torch_gpu.zeros_like(s)
    fs.solveFPS(p, s)
    po = O.fft_solve(go, so)
    e = rel(to_host(p), po)
    F.makeVelocityDivFreeFast(g, adt, rho0, d["rho"], p, d["ux"], d["uy"], d["uz"], d["psi1"], d["psi2"], dst1, dst2)
    uo = O.project_fast(go, adt, rho0, rho, po, psi1, psi2, ux, uy, uz, st1, st2)
    eu = max(rel(to_host(a), b) for a, b in zip((d["ux"], d["uy"], d["uz"]), uo))
    md = F.max_divergence(g, d["ux"], d["uy"], d["uz"])
    scale = np.abs(O.divergence(go, ux, uy, uz)).max()
    print("  pressure %.2e, velocities %.2e, max div %.2e of %.2e" % (e, eu, md, scale), flush=True)
    assert e < 1e-11 and eu < 1e-11
    assert md <= 1e-10 * scale and md <= 1e-10
    fs.close()


def test_update_pressure_kernel(torch_gpu):
    """updatePressure + clean_st_prev_bc (poissonEqn.f90:374-459) on the device against
    the oracle: p = psi, both rotations, old st levels zero on the wall faces only."""
    n = (48, 20, 12)
    print("case update pressure", n, flush=True)
    g, go = grids(n, 1)
    psi, psi1, psi2 = (rng.standard_normal(n) for _ in range(3))
    st, st1, st2 = random_st(n), random_st(n), random_st(n)
    dv = lambda v: tuple(to_dev(a, torch_gpu) for a in v)  # noqa: E731
    dpsi, d1, d2 = (to_dev(a, torch_gpu) for a in (psi, psi1, psi2))
    dp = torch_gpu.zeros_like(dpsi)
    dst, dst1, dst2 = dv(st), dv(st1), dv(st2)
    F.updatePressure(g, dpsi, dp, d1, d2, dst, dst1, dst2)
    p, n1, n2, s1, s2 = O.update_pressure(go, psi, psi1, psi2, st, st1, st2)
    for got, want in ((dp, p), (d1, n1), (d2, n2)):
        assert np.array_equal(to_host(got), want)
    for got, want in zip(dst1 + dst2, s1 + s2):
        assert np.array_equal(to_host(got), want)
    # without surface tension levels only the pressure levels move
    F.updatePressure(g, dpsi, None, d1, d2)
    assert np.array_equal(to_host(d2), psi) and np.array_equal(to_host(d1), psi)


def test_poisson_eqn_fast_steps_rotate_old_levels(torch_gpu):
    """Four FAST_MODE pressure steps through PoissonEqn (source with the old levels
    psi^n, psi^(n-1), st^n, st^(n-1), solveFPS, projection, updatePressure with the
    surface-tension rotation and wall clean-up), started from given old levels as
    poissonEqnCTOR and the VOF constructor set them, against the oracle doing the
    same sequence."""
    n = (32, 24, 16)
    print("case PoissonEqn FAST steps", n, flush=True)
    g, go = grids(n, 1)
    rho, _ = F.make_bubble_density(g, 3, 0.3, 1004, nref=6)
    rho_l, rho_g, adt = 1.0, 1.0 / 20.0, 1e-3
    eq = T.PoissonEqn(g, "FAST", rho_l, rho_g)
    drho = to_dev(rho, torch_gpu)
    o1, o2 = rng.standard_normal(n), rng.standard_normal(n)
    t1, t2 = random_st(n), random_st(n)   # as read from the restart files: walls not cleaned yet
    dv = lambda v: tuple(to_dev(a, torch_gpu) for a in v)  # noqa: E731
    eq.setOldLevels(to_dev(o1, torch_gpu), to_dev(o2, torch_gpu), dv(t1), dv(t2))
    psi = to_dev(o1, torch_gpu)
    p = torch_gpu.zeros_like(drho)
    for step in range(4):
        ux, uy, uz = F.make_velocity(g, 2000 + step)
        dux, duy, duz = (to_dev(a, torch_gpu) for a in (ux, uy, uz))
        eq.solvePoissonEqn(psi, drho, dux, duy, duz, adt)
        eq.makeVelocityDivFree(psi, drho, dux, duy, duz, adt)
        so = O.source_fast(go, adt, rho_g, rho, ux, uy, uz, o1, o2, t1, t2)
        po = O.fft_solve(go, so)
        uxo, uyo, uzo = O.project_fast(go, adt, rho_g, rho, po, o1, o2, ux, uy, uz, t1, t2)
        assert rel(to_host(psi), po) < 1e-12, step
        assert rel(to_host(dux), uxo) < 1e-12 and rel(to_host(duy), uyo) < 1e-12, step
        # step 0 reads st levels whose wall faces are not cleaned yet: interior rows only
        assert max_div(go, dux, duy, duz, walls_clean=step > 0) <= 1e-10 * np.abs(O.divergence(go, ux, uy, uz)).max(), step
        st = random_st(n)   # this step's surface tension (computeSurfaceTension, out of this path)
        eq.updatePressure(psi, p, dv(st))
        _, o1, o2, t1, t2 = O.update_pressure(go, po, o1, o2, st, t1, t2)
        assert np.array_equal(to_host(p), to_host(psi))
        for got, want in zip(eq.st_old_[0] + eq.st_old_[1], t1 + t2):
            assert np.array_equal(to_host(got), want), step
    eq.close()
