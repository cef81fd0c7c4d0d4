"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned by what the mathematics fixes: closed-form coordinates and
eigenvalues, numpy's FFT, dense linear-algebra solves on tiny grids, symmetry,
null space and conservation invariants, and exactness of the transfer
operators on linear functions.  Each check is chosen so that a plausible
mistake (wrong index or metric, sign, face average, colour order) fails it.
"""
import zlib

import numpy as np
import pytest

from oracle import pyoracle as O

TWO_PI = 2 * np.pi


def grid(n, stretch_y=0, gamma=1.4, L=(TWO_PI, 2.0, np.pi)):
    return O.desc(n, L=L, stretch=(0, stretch_y, 0), gamma=gamma)


def tanh_c(i, n, L, g):
    return 0.5 * L * (1 + np.tanh(-g + 2 * g * (i - 0.5) / n) / np.tanh(g))


def tanh_f(i, n, L, g):
    return 0.5 * L * (1 + np.tanh(-g + 2 * g * i / n) / np.tanh(g))


def centres(g, a):
    n = g.n[a]
    xc, _ = O.axis_coords(n, g.L[a], g.stretch[a], g.gamma, 0)
    return xc


def faces(g, a):
    n = g.n[a]
    _, xf = O.axis_coords(n, g.L[a], g.stretch[a], g.gamma, 0)
    return xf  # f(0..n)


def vol(g):
    dx = np.diff(faces(g, 0))
    dy = np.diff(faces(g, 1))
    dz = np.diff(faces(g, 2))
    return dx[:, None, None] * dy[None, :, None] * dz[None, None, :]


def dense(fn, shape):
    n = int(np.prod(shape))
    M = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        M[:, c] = fn(e.reshape(shape, order="F")).ravel(order="F")
    return M


rng = np.random.default_rng(1234)


@pytest.fixture(autouse=True)
def _rng_per_test(request):
    """Reseeds the shared generator from the test's name, so that each test draws the same
    values whatever ran before it (alone, in file order or split across workers)."""
    global rng
    rng = np.random.default_rng(zlib.crc32(request.node.name.encode()))


# ------------------------------------------------------------ coordinates (a30)

@pytest.mark.parametrize("stretch,gamma", [(0, 1.4), (1, 1.4), (1, 1.5)])
def test_axis_coordinates_closed_form(stretch, gamma):
    n, L, G = 24, 2.0, 2
    xc, xf = O.axis_coords(n, L, stretch, gamma, G)
    i_c = np.arange(1 - G, n + G + 1)
    i_f = np.arange(-G, n + G + 1)
    if stretch:
        ec, ef = tanh_c(i_c, n, L, gamma), tanh_f(i_f, n, L, gamma)
    else:
        d = L / n
        ec, ef = (i_c - 0.5) * d, i_f * d
    assert np.allclose(xc, ec, rtol=0, atol=1e-15)
    assert np.allclose(xf, ef, rtol=0, atol=1e-15)
    # walls at faces 0 and n, coarse centres coincide with fine faces 2I-1
    assert abs(xf[G]) < 1e-15 and abs(xf[G + n] - L) < 1e-14
    xc2, _ = O.axis_coords(n // 2, L, stretch, gamma, 0)
    assert np.allclose(xc2, xf[G + 1:G + n:2], atol=1e-15)


# ------------------------------------------------------------------ FFT (a2)

@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 32, 64, 96, 128, 192, 256])
def test_fft_matches_numpy(n):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(O.fft_c2c(x, -1), np.fft.fft(x), atol=1e-12 * n)
    assert np.allclose(O.fft_c2c(x, +1), np.fft.ifft(x) * n, atol=1e-12 * n)


# ------------------------------------------------------- operator (a16, a24)

def test_operator_eigenmodes_uniform():
    g = grid((16, 12, 8))
    nx, ny, nz = g.shape if hasattr(g, "shape") else O.shape(g)
    xc, yc, zc = centres(g, 0), centres(g, 1), centres(g, 2)
    dx, dy, dz = TWO_PI / nx, 2.0 / ny, np.pi / nz
    beta = np.ones((nx, ny, nz), order="F")
    for m in (1, 3):
        v = np.cos(TWO_PI * m * xc / TWO_PI)[:, None, None] * np.ones((1, ny, nz))
        lam = 2 * (np.cos(TWO_PI * m / nx) - 1) / dx ** 2
        assert np.allclose(O.apply_op(g, beta, v), lam * v, atol=1e-11)
        v = np.ones((nx, 1, 1)) * np.cos(np.pi * m * yc / 2.0)[None, :, None] * np.ones((1, 1, nz))
        lam = 2 * (np.cos(np.pi * m / ny) - 1) / dy ** 2
        assert np.allclose(O.apply_op(g, beta, v), lam * v, atol=1e-11)
        v = np.ones((nx, ny, 1)) * np.cos(TWO_PI * m * zc / np.pi)[None, None, :]
        lam = 2 * (np.cos(TWO_PI * m / nz) - 1) / dz ** 2
        assert np.allclose(O.apply_op(g, beta, v), lam * v, atol=1e-10)


def test_operator_quadratic_on_tanh_walls():
    g = grid((4, 20, 4), stretch_y=1, gamma=1.5)
    ny = 20
    yc = tanh_c(np.arange(1, ny + 1), ny, 2.0, 1.5)
    yf = tanh_f(np.arange(0, ny + 1), ny, 2.0, 1.5)
    p = np.ones((4, 1, 4)) * (yc ** 2)[None, :, None]
    q = O.apply_op(g, np.ones((4, ny, 4), order="F"), p)[0, :, 0]
    dyf = np.diff(yf)
    expect = np.empty(ny)
    expect[1:-1] = (yc[2:] - yc[:-2]) / dyf[1:-1]       # flux (y_{j+1}+y_j) differences
    expect[0] = (yc[1] + yc[0]) / dyf[0]                # no flux through the wall
    expect[-1] = -(yc[-1] + yc[-2]) / dyf[-1]
    assert np.allclose(q, expect, rtol=1e-12, atol=1e-10)


def test_operator_arithmetic_face_average():
    g = grid((16, 4, 4))
    nx = 16
    dx = TWO_PI / nx
    beta = np.ones((nx, 4, 4), order="F")
    beta[8:] = 2.0
    xc = centres(g, 0)
    p = xc[:, None, None] * np.ones((1, 4, 4))
    q = O.apply_op(g, beta, p)
    assert np.allclose(q[7], (1.5 - 1.0) / dx, atol=1e-11)   # interface cell, left side
    assert np.allclose(q[8], (2.0 - 1.5) / dx, atol=1e-11)
    assert np.allclose(q[3], 0.0, atol=1e-11)


@pytest.mark.parametrize("stretch", [0, 1])
def test_operator_symmetry_nullspace_conservation(stretch):
    g = grid((6, 8, 4), stretch_y=stretch)
    shape = O.shape(g)
    beta = np.exp(rng.standard_normal(shape))
    V = vol(g)
    A = dense(lambda e: O.apply_op(g, beta, e), shape)
    W = np.diag(V.ravel(order="F"))
    VA = W @ A
    assert np.allclose(VA, VA.T, atol=1e-10 * np.abs(VA).max())
    assert np.allclose(A @ np.ones(A.shape[0]), 0, atol=1e-10 * np.abs(A).max())
    ev = np.linalg.eigvalsh(0.5 * (VA + VA.T))
    assert ev.max() < 1e-9 * np.abs(ev).max() and np.sum(np.abs(ev) < 1e-9 * np.abs(ev).max()) == 1
    u = rng.standard_normal(shape)
    assert abs(np.sum(V * O.apply_op(g, beta, u))) < 1e-10 * np.abs(VA).max()


# --------------------------------------------------------------- FFT path (a3-a7)

def test_fft_operator_equals_unit_beta_operator():
    for stretch in (0, 1):
        g = grid((8, 10, 6), stretch_y=stretch)
        p = rng.standard_normal(O.shape(g))
        q1 = O.fft_operator(g, p)
        q2 = O.apply_op(g, np.ones(O.shape(g)), p)
        assert np.allclose(q1, q2, rtol=1e-12, atol=1e-10)


def test_tridiag_vs_dense():
    g = grid((4, 24, 4), stretch_y=1, gamma=1.5)
    ny = 24
    yc = tanh_c(np.arange(0, ny + 2), ny, 2.0, 1.5)   # c(0..ny+1)
    yf = tanh_f(np.arange(0, ny + 1), ny, 2.0, 1.5)
    dyc = np.diff(yc)          # dyc(j) for j=1..ny+1
    dyf = np.diff(yf)          # dyf(j) for j=1..ny
    am = 1 / (dyc[:ny] * dyf)
    ap = 1 / (dyc[1:] * dyf)
    am[0] = 0
    ap[-1] = 0
    d = rng.standard_normal(ny) + 1j * rng.standard_normal(ny)
    for lam in (-7.0, -1e-4):
        M = np.diag(lam - am - ap) + np.diag(am[1:], -1) + np.diag(ap[:-1], 1)
        x = O.tridiag(g, lam, 0, d)
        # backward error (the lam -> 0 system has condition ~1e7)
        assert np.linalg.norm(M @ x - d) < 1e-13 * np.abs(M).max() * np.linalg.norm(x)
        if lam < -1:
            assert np.allclose(x, np.linalg.solve(M, d), rtol=1e-12, atol=1e-14)
    M = np.diag(-am - ap) + np.diag(am[1:], -1) + np.diag(ap[:-1], 1)
    x = O.tridiag(g, 0.0, 1, d)
    assert x[-1] == 0
    assert np.allclose((M @ x)[:-1], d[:-1], atol=1e-9)


@pytest.mark.parametrize("stretch", [0, 1])
def test_fft_solve_dense_tiny(stretch):
    g = grid((8, 6, 4), stretch_y=stretch)
    shape = O.shape(g)
    V = vol(g)
    s = rng.standard_normal(shape)
    s -= np.sum(V * s) / V.sum()              # compatible right-hand side
    p = O.fft_solve(g, s)
    L = dense(lambda e: O.fft_operator(g, e), shape)
    n = L.shape[0]
    aug = np.zeros((n + 1, n + 1))
    aug[:n, :n] = L
    aug[:n, n] = V.ravel(order="F")
    aug[n, :n] = V.ravel(order="F")
    ref = np.linalg.solve(aug, np.r_[s.ravel(order="F"), 0.0])[:n].reshape(shape, order="F")
    assert np.allclose(p, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    assert abs(O.volume_mean(g, p)) < 1e-14 * np.abs(p).max()


@pytest.mark.parametrize("n,stretch", [((64, 64, 64), 0), ((48, 32, 24), 1)])
def test_fft_solve_residual(n, stretch):
    g = grid(n, stretch_y=stretch, gamma=1.5)
    V = vol(g)
    s = rng.standard_normal(O.shape(g))
    s -= np.sum(V * s) / V.sum()
    p = O.fft_solve(g, s)
    r = O.fft_operator(g, p) - s
    assert np.linalg.norm(r) < 1e-12 * np.linalg.norm(s)
    assert abs(O.volume_mean(g, p)) < 1e-12 * np.abs(p).max()


# -------------------------------------------------------- transfer ops (a27-a28)

@pytest.mark.parametrize("stretch", [0, 1])
def test_restriction_exact_on_linear(stretch):
    g = grid((8, 16, 8), stretch_y=stretch, gamma=1.5)
    xc, yc, zc = centres(g, 0), centres(g, 1), centres(g, 2)
    f = 1.5 + 0.3 * xc[:, None, None] - 0.7 * yc[None, :, None] + 0.2 * zc[None, None, :]
    c = O.restrict(g, 0, f)
    gc = grid((4, 8, 4), stretch_y=stretch, gamma=1.5)
    Xc, Yc, Zc = centres(gc, 0), centres(gc, 1), centres(gc, 2)
    expect = 1.5 + 0.3 * Xc[:, None, None] - 0.7 * Yc[None, :, None] + 0.2 * Zc[None, None, :]
    assert np.allclose(c, expect, atol=1e-13)
    if not stretch:   # uniform: plain average of the 8 children
        avg = f.reshape(2, 4, 2, 8, 2, 4, order="F").mean(axis=(0, 2, 4))
        assert np.allclose(c, avg, atol=1e-14)


@pytest.mark.parametrize("stretch", [0, 1])
def test_prolongation(stretch):
    g = grid((8, 16, 8), stretch_y=stretch, gamma=1.5)
    cs = O.shape(g, 1)
    fine0 = rng.standard_normal(O.shape(g))
    out = O.prolong_add(g, 0, 2.5 * np.ones(cs), fine0)
    assert np.allclose(out - fine0, 2.5, atol=1e-13)          # constants, walls and wrap included
    gc = grid((4, 8, 4), stretch_y=stretch, gamma=1.5)
    Yc, yc = centres(gc, 1), centres(g, 1)
    lin = np.ones((4, 1, 4)) * Yc[None, :, None]
    out = O.prolong_add(g, 0, lin, np.zeros(O.shape(g)))
    inner = slice(1, 15)                                       # away from the wall ghosts
    assert np.allclose(out[:, inner, :], np.ones((8, 1, 8)) * yc[None, inner, None], atol=1e-13)
    if not stretch:
        delta = np.zeros(cs)
        delta[1, 3, 1] = 1.0
        out = O.prolong_add(g, 0, delta, np.zeros(O.shape(g)))
        vals = sorted(set(np.round(out[out > 0] * 64, 10)))
        assert vals == [1.0, 3.0, 9.0, 27.0]
        assert np.isclose(out.sum(), 8.0)


# --------------------------------------------------------- smoother (a22-a25)

def test_rbgs_ordering_and_fixed_point():
    g = grid((8, 8, 8), stretch_y=1)
    shape = O.shape(g)
    beta = np.exp(0.5 * rng.standard_normal(shape))
    V = vol(g)
    q = rng.standard_normal(shape)
    q -= np.sum(V * q) / V.sum()           # compatible: sum V q = 0
    p0 = rng.standard_normal(shape)
    p1 = O.rbgs(g, 0, beta, q, p0, 1, remove_mean=0)
    r = O.residual(g, 0, beta, q, p1)
    i, j, k = np.meshgrid(np.arange(1, 9), np.arange(1, 9), np.arange(1, 9), indexing="ij")
    black = ((i + j + k) % 2) == 1
    interior = (j > 1) & (j < 8)
    # black cells were updated last with final red neighbours: exact there
    assert np.abs(r[black & interior]).max() < 1e-12 * np.abs(q).max()
    # red cells see new black values: generally not exact
    assert np.abs(r[~black & interior]).max() > 1e-6
    # at a wall the ghost is the cell's own value from before its colour's
    # half-sweep (rbgs.f90:202-230 reads the halo the previous updateBoundaries
    # filled), so the wall face stays in the update: a black wall cell solves
    # (sum_int a + a_w) p1 = sum_int a p_nb + a_w p0 - q, while A (exact Neumann
    # after the fill) has no wall flux.  Its residual is therefore
    # q - A p1 = -a_w (p1 - p0) with a_w = beta/(dyc_wall dyf): non-zero, where an
    # exact-Neumann smoother would leave zero.
    yc, yf = O.axis_coords(8, g.L[1], g.stretch[1], g.gamma, 1)   # yc[i] = centre i, yf[1 + j] = face j
    a_bot = beta[:, 0, :] / ((yc[1] - yc[0]) * (yf[2] - yf[1]))
    a_top = beta[:, 7, :] / ((yc[9] - yc[8]) * (yf[9] - yf[8]))
    for jj, aw in ((0, a_bot), (7, a_top)):
        m = black[:, jj, :]
        want = -aw * (p1[:, jj, :] - p0[:, jj, :])
        assert np.abs(want[m]).min() > 1e-6
        assert np.allclose(r[:, jj, :][m], want[m], rtol=1e-11, atol=1e-12 * np.abs(q).max())
    A = dense(lambda e: O.apply_op(g, beta, e), shape)
    sol = np.linalg.lstsq(A, q.ravel(order="F"), rcond=None)[0].reshape(shape, order="F")
    p2 = O.rbgs(g, 0, beta, q, sol, 3, remove_mean=1)
    assert np.allclose(p2, sol - sol.mean(), atol=1e-10 * np.abs(sol).max())


# ------------------------------------------------------- coarse solve (a26)

def test_coarse_exact_dense():
    g = grid((4, 6, 4), stretch_y=1, gamma=1.5)
    shape = O.shape(g)
    beta = np.exp(rng.standard_normal(shape))
    q = rng.standard_normal(shape)
    V = vol(g).ravel(order="F")
    A = dense(lambda e: O.apply_op(g, beta, e), shape)
    M = -V[:, None] * A
    n = M.shape[0]
    aug = np.zeros((n + 1, n + 1))
    aug[:n, :n] = M
    aug[:n, n] = 1.0
    aug[n, :n] = 1.0
    ref = np.linalg.solve(aug, np.r_[-V * q.ravel(order="F"), 0.0])[:n].reshape(shape, order="F")
    got = O.coarse_exact(g, 0, beta, q)
    assert np.allclose(got, ref, atol=1e-11 * np.abs(ref).max())


# ------------------------------------------------------------ PCG (a15, a21)

@pytest.mark.parametrize("stretch", [0, 1])
def test_pcg_dense_tiny(stretch):
    # Own generator: the draw must not depend on which tests ran before.  On some draws the
    # non-flexible PCG stagnates near 5e-12 on the tanh grid (the red-black V-cycle is not a
    # symmetric preconditioner), so the case is fixed to one that reaches 1e-13 in every mode.
    rng = np.random.default_rng(1)
    g = grid((8, 8, 8), stretch_y=stretch)
    shape = O.shape(g)
    beta = np.exp(rng.standard_normal(shape))
    V = vol(g)
    q = rng.standard_normal(shape)
    q -= np.sum(V * q) / V.sum()
    A = dense(lambda e: O.apply_op(g, beta, e), shape)
    ref = np.linalg.lstsq(A, q.ravel(order="F"), rcond=None)[0].reshape(shape, order="F")
    ref -= ref.mean()
    for coarse_mode, flexible in ((0, 0), (1, 0), (1, 1)):
        prm = O.params(n_levels=3, tol=1e-13, max_iter=200, coarse_mode=coarse_mode, coarse_cycles=6,
                       flexible=flexible)
        psi, it, res, rc = O.pcg_solve(g, prm, beta, q)
        assert rc == 0 and res <= 1e-13
        r = q - O.apply_op(g, beta, psi)
        assert abs(np.linalg.norm(r) / np.linalg.norm(q) - res) < 1e-14   # recursive vs true residual
        assert np.allclose(psi - psi.mean(), ref, atol=1e-10 * np.abs(ref).max())


def test_pcg_max_iter_and_warm_start():
    g = grid((16, 16, 16))
    shape = O.shape(g)
    beta = np.ones(shape)
    q = rng.standard_normal(shape)
    q -= q.mean()
    psi, it, res, rc = O.pcg_solve(g, O.params(n_levels=3, tol=1e-14, max_iter=2), beta, q)
    assert rc == 1 and it == 2
    psi2, it2, res2, rc2 = O.pcg_solve(g, O.params(n_levels=3, tol=1e-8, max_iter=50), beta, q, psi0=psi)
    assert rc2 == 0 and res2 <= 1e-8
    psi3, it3, _, _ = O.pcg_solve(g, O.params(n_levels=3, tol=1e-8, max_iter=50), beta, q)
    assert it2 < it3


@pytest.mark.parametrize("stretch", [0, 1])
def test_mg_tolerance_loop(stretch):
    """solveMG's loop (mg_tol > 0): the cycle count is the first c at which the
    fixed c-cycle application's residual ||r - A z|| / ||r|| is at most tolMG,
    capped at maxIterMG, and the result is that application's, bit for bit;
    before any cycle the residual is 1, so there is always at least one."""
    g = grid((32, 16, 32), stretch_y=stretch)
    shape = O.shape(g)
    beta = np.exp(0.8 * rng.standard_normal(shape))
    r = rng.standard_normal(shape)
    r -= np.sum(vol(g) * r) / vol(g).sum()  # in the range of A (sum V r = 0)
    fixed, res = {}, {}
    for c in range(1, 7):
        fixed[c] = O.mg_apply(g, O.params(n_levels=3, mg_cycles=c), beta, r)
        res[c] = np.linalg.norm(r - O.apply_op(g, beta, fixed[c])) / np.linalg.norm(r)
    assert all(res[c + 1] < res[c] for c in range(1, 6))
    for tol in (0.9, res[2] * 1.0001, res[3] * 0.9999, res[5] * 1.0001):
        want = next(c for c in range(1, 7) if res[c] <= tol)
        z = O.mg_apply(g, O.params(n_levels=3, mg_tol=tol, mg_max_iter=100), beta, r)
        assert O.last_mg_cycles() == (want, want), (tol, want)
        assert np.array_equal(z, fixed[want])
    for cap in (1, 2, 4):
        z = O.mg_apply(g, O.params(n_levels=3, mg_tol=1e-300, mg_max_iter=cap), beta, r)
        assert O.last_mg_cycles() == (cap, cap) and np.array_equal(z, fixed[cap])
    # in a solve: every application counted, the converged answer unchanged
    q = r
    psi, it, rres, rc = O.pcg_solve(g, O.params(n_levels=3, tol=1e-10, max_iter=200, mg_tol=0.05,
                                                flexible=1), beta, q)
    total, most, last = O.last_mg_cycles(with_last=True)
    assert rc == 0 and it + 1 <= total <= (it + 1) * most and most >= 2 and 1 <= last <= most
    psi1, _, _, rc1 = O.pcg_solve(g, O.params(n_levels=3, tol=1e-10, max_iter=200, flexible=1), beta, q)
    assert np.linalg.norm((psi - psi.mean()) - (psi1 - psi1.mean())) <= 1e-8 * np.linalg.norm(psi1 - psi1.mean())


# ---------------------------------------------- projection and divergence (a33)

def test_divergence_manufactured():
    g = grid((16, 8, 8), stretch_y=1, gamma=1.5)
    nx, ny, nz = O.shape(g)
    xf, yf, zf = faces(g, 0), faces(g, 1), faces(g, 2)
    ux = np.sin(TWO_PI * xf[1:] / TWO_PI)[:, None, None] * np.ones((1, ny, nz))
    uy = np.ones((nx, 1, nz)) * (yf ** 2)[None, :, None]
    uz = np.ones((nx, ny, 1)) * np.cos(TWO_PI * zf[1:] / np.pi)[None, None, :]
    d = O.divergence(g, ux, uy, uz)
    ex = (np.sin(xf[1:]) - np.sin(xf[:-1])) / np.diff(xf)
    ey = (yf[1:] ** 2 - yf[:-1] ** 2) / np.diff(yf)
    ez = (np.cos(2 * zf[1:]) - np.cos(2 * zf[:-1])) / np.diff(zf)
    expect = ex[:, None, None] + ey[None, :, None] + ez[None, None, :]
    assert np.allclose(d, expect, atol=1e-12)


def test_projection_identity_mg():
    g = grid((16, 16, 16), stretch_y=1, gamma=1.5)
    shape = O.shape(g)
    nx, ny, nz = shape
    ux, uz = rng.standard_normal(shape), rng.standard_normal(shape)
    uy = rng.standard_normal((nx, ny + 1, nz))
    uy[:, 0, :] = 0
    uy[:, -1, :] = 0
    rho = np.where(rng.random(shape) < 0.2, 1 / 20, 1.0)
    adt = 1e-3
    q = O.divergence(g, ux, uy, uz)
    beta = O.beta_from_rho(g, adt, rho)
    # a cell-by-cell random density with contrast 20 makes the reference's
    # Fletcher-Reeves PCG with its red/black V-cycle stagnate; the identity
    # below does not depend on the CG variant, so the flexible one is used
    psi, it, res, rc = O.pcg_solve(g, O.params(n_levels=3, tol=1e-12, max_iter=300, flexible=1), beta, q)
    assert rc == 0
    ux2, uy2, uz2 = O.project_mg(g, adt, rho, psi, ux, uy, uz)
    div = O.divergence(g, ux2, uy2, uz2)
    r = q - O.apply_op(g, beta, psi)
    assert np.allclose(div, r, atol=1e-12 * np.abs(q).max())     # div u_new = q - A psi
    assert np.abs(div).max() < 1e-10 * np.abs(q).max()
    assert np.all(uy2[:, 0, :] == 0) and np.all(uy2[:, -1, :] == 0)


def test_fast_source_and_projection_constant_density():
    g = grid((32, 32, 32))
    shape = O.shape(g)
    nx, ny, nz = shape
    ux, uz = rng.standard_normal(shape), rng.standard_normal(shape)
    uy = rng.standard_normal((nx, ny + 1, nz))
    uy[:, 0, :] = 0
    uy[:, -1, :] = 0
    rho = np.ones(shape)
    z = np.zeros(shape)
    dt = 1e-3
    s = O.source_fast(g, dt, 1.0, rho, ux, uy, uz, z, z)
    assert np.allclose(s, O.divergence(g, ux, uy, uz) / dt, rtol=1e-13, atol=1e-9)
    p = O.fft_solve(g, s)
    ux2, uy2, uz2 = O.project_fast(g, dt, 1.0, rho, p, z, z, ux, uy, uz)
    d = O.divergence(g, ux2, uy2, uz2)
    assert np.abs(d).max() < 1e-10


def test_fast_source_variable_density_is_conservative():
    g = grid((16, 16, 8), stretch_y=1)
    shape = O.shape(g)
    nx, ny, nz = shape
    z3 = np.zeros(shape)
    zy = np.zeros((nx, ny + 1, nz))
    rho = np.where(rng.random(shape) < 0.3, 1 / 20, 1.0)
    psi1, psi2 = rng.standard_normal(shape), rng.standard_normal(shape)
    s = O.source_fast(g, 1e-3, 1 / 20, rho, z3, zy, z3, psi1, psi2)
    assert np.abs(s).max() > 0
    assert abs(np.sum(vol(g) * s)) < 1e-11 * np.abs(vol(g) * s).sum()
    s0 = O.source_fast(g, 1e-3, 1.0, np.ones(shape), z3, zy, z3, psi1, psi2)
    assert np.abs(s0).max() == 0


def _random_u(shape, walls_zero=True):
    nx, ny, nz = shape
    ux, uz = rng.standard_normal(shape), rng.standard_normal(shape)
    uy = rng.standard_normal((nx, ny + 1, nz))
    if walls_zero:
        uy[:, 0, :] = 0
        uy[:, -1, :] = 0
    return ux, uy, uz


def _face_gradient(g, phi):
    """The staggered discrete gradient of a cell field, laid out like the velocity
    (x, z: the face after each cell, periodic; y: faces 0..ny with the walls, where
    the Neumann ghost makes it zero): (phi_nb - phi_c) / (centre distance)."""
    nx, ny, nz = O.shape(g)
    xc = centres(g, 0)
    yc, _ = O.axis_coords(ny, g.L[1], g.stretch[1], g.gamma, 0)
    zc = centres(g, 2)
    gx = (np.roll(phi, -1, axis=0) - phi) / (xc[1] - xc[0])
    gz = (np.roll(phi, -1, axis=2) - phi) / (zc[1] - zc[0])
    gy = np.zeros((nx, ny + 1, nz))
    gy[:, 1:ny, :] = np.diff(phi, axis=1) / np.diff(yc)[None, :, None]
    return gx, gy, gz


def _vmean(g, a):
    V = vol(g)
    return a - np.sum(V * a) / V.sum()


@pytest.mark.parametrize("stretch", [0, 1])
def test_fast_split_fixed_point_variable_density(stretch):
    """The splitting's fixed point (Dodd-Ferrante / Cifani, poissonEqn.f90:193-305,
    momentumEqn.f90:999-1127): if the extrapolated old pressure 2 psi^n - psi^(n-1)
    is the exact variable-density pressure psi_e (div((1/rho)_f grad psi_e) =
    div u* / (a dt), the MG_MODE equation with beta = a dt / rho, solved densely),
    the source is L psi_e, the constant-coefficient solve returns psi_e and the
    projection equals the variable-density projection with psi_e.  psi^n and
    psi^(n-1) differ (psi_e + phi, psi_e + 2 phi), so the result pins the sign of
    the old-pressure term and the 2, -1 weights: a flipped sign, a dropped
    psi^(n-1) or psi^n alone, another face mean of 1/rho or another rho0 all fail."""
    g = grid((8, 8, 4), stretch_y=stretch, gamma=1.5)
    shape = O.shape(g)
    rho_l, rho_g = 1.0, 1.0 / 20.0
    rho0 = min(rho_l, rho_g)
    rho = np.where(rng.random(shape) < 0.35, rho_g, rho_l)
    adt = 2e-3
    ux, uy, uz = _random_u(shape)
    q = O.divergence(g, ux, uy, uz)
    beta = O.beta_from_rho(g, adt, rho)
    A = dense(lambda e: O.apply_op(g, beta, e), shape)
    psi_e = np.linalg.lstsq(A, q.ravel(order="F"), rcond=None)[0].reshape(shape, order="F")
    assert np.abs(q - O.apply_op(g, beta, psi_e)).max() < 1e-10 * np.abs(q).max()
    phi = rng.standard_normal(shape)
    psi1, psi2 = psi_e + phi, psi_e + 2 * phi
    s = O.source_fast(g, adt, rho0, rho, ux, uy, uz, psi1, psi2)
    p = O.fft_solve(g, s)
    ref = _vmean(g, psi_e)
    assert np.abs(p - ref).max() < 1e-10 * np.abs(ref).max()
    fast = O.project_fast(g, adt, rho0, rho, p, psi1, psi2, ux, uy, uz)
    mg = O.project_mg(g, adt, rho, psi_e, ux, uy, uz)
    for a, b in zip(fast, mg):
        assert np.abs(a - b).max() < 1e-10 * np.abs(b).max()
    # the two old levels enter only through 2 psi^n - psi^(n-1): with those swapped
    # the fixed point is lost (the test is sensitive to the weights)
    s_sw = O.source_fast(g, adt, rho0, rho, ux, uy, uz, psi2, psi1)
    assert np.abs(O.fft_solve(g, s_sw) - ref).max() > 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize("stretch", [0, 1])
def test_fast_surface_tension_is_a_pressure_gauge(stretch):
    """The old surface-tension levels enter the face value as
    G_f = d_f(2 psi^n - psi^(n-1)) - 2 st^n_f + st^(n-1)_f (poissonEqn.f90:255-301,
    momentumEqn.f90:1104-1121).  So st levels that are discrete gradients of cell
    fields phi1, phi2 act exactly like the pressure levels psi^n - phi1 and
    psi^(n-1) - phi2 with no surface tension.  A flipped sign (psi + phi), a
    dropped level, other weights or an st face off by one cell all fail."""
    g = grid((16, 8, 8), stretch_y=stretch, gamma=1.5)
    shape = O.shape(g)
    rho = np.where(rng.random(shape) < 0.3, 1 / 20, 1.0)
    adt = 1e-3
    ux, uy, uz = _random_u(shape)
    psi1, psi2 = rng.standard_normal(shape), rng.standard_normal(shape)
    phi1, phi2 = rng.standard_normal(shape), rng.standard_normal(shape)
    st1, st2 = _face_gradient(g, phi1), _face_gradient(g, phi2)
    s = O.source_fast(g, adt, 1 / 20, rho, ux, uy, uz, psi1, psi2, st1, st2)
    s_ref = O.source_fast(g, adt, 1 / 20, rho, ux, uy, uz, psi1 - phi1, psi2 - phi2)
    assert np.abs(s - s_ref).max() < 1e-12 * np.abs(s_ref).max()
    psi = rng.standard_normal(shape)
    u = O.project_fast(g, adt, 1 / 20, rho, psi, psi1, psi2, ux, uy, uz, st1, st2)
    u_ref = O.project_fast(g, adt, 1 / 20, rho, psi, psi1 - phi1, psi2 - phi2, ux, uy, uz)
    for a, b in zip(u, u_ref):
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
    # the st terms are not inert: the same levels with the opposite sign differ
    s_neg = O.source_fast(g, adt, 1 / 20, rho, ux, uy, uz, psi1 + phi1, psi2 + phi2)
    assert np.abs(s - s_neg).max() > 1e-2 * np.abs(s_ref).max()


def test_fast_step_divergence_free_with_history():
    """One FAST_MODE step with everything present (two-valued density, distinct
    psi^n, psi^(n-1), surface-tension levels that are not gradients): source, FFT
    solve, projection.  The source and the projection use the same G_f, so div u
    vanishes to round-off whatever G_f is; a term in one but not the other fails."""
    g = grid((32, 24, 16), stretch_y=1, gamma=1.5)
    shape = O.shape(g)
    nx, ny, nz = shape
    rho = np.where(rng.random(shape) < 0.3, 1 / 20, 1.0)
    adt = 1e-3
    ux, uy, uz = _random_u(shape)
    psi1, psi2 = rng.standard_normal(shape), rng.standard_normal(shape)
    st1 = _random_u(shape)
    st2 = _random_u(shape)
    s = O.source_fast(g, adt, 1 / 20, rho, ux, uy, uz, psi1, psi2, st1, st2)
    p = O.fft_solve(g, s)
    ux2, uy2, uz2 = O.project_fast(g, adt, 1 / 20, rho, p, psi1, psi2, ux, uy, uz, st1, st2)
    d = O.divergence(g, ux2, uy2, uz2)
    scale = np.abs(O.divergence(g, ux, uy, uz)).max()
    assert np.abs(d).max() < 1e-10 * scale
    # without the st levels in the projection the step is not divergence free
    ux3, uy3, uz3 = O.project_fast(g, adt, 1 / 20, rho, p, psi1, psi2, ux, uy, uz)
    assert np.abs(O.divergence(g, ux3, uy3, uz3)).max() > 1e-3 * scale


def test_update_pressure_rotates_and_cleans_walls():
    """updatePressure (poissonEqn.f90:374-415): p = psi, psi^(n-1) <- psi^n <- psi,
    st^(n-1) <- st^n <- st, then clean_st_prev_bc (:419-459) zeroes both old st
    levels on the Neumann wall faces (y faces 0 and ny) and nowhere else."""
    g = grid((8, 6, 4))
    shape = O.shape(g)
    nx, ny, nz = shape
    psi, psi1, psi2 = (rng.standard_normal(shape) for _ in range(3))
    st = _random_u(shape, walls_zero=False)
    st1 = _random_u(shape, walls_zero=False)
    st2 = _random_u(shape, walls_zero=False)
    p, n1, n2, s1, s2 = O.update_pressure(g, psi, psi1, psi2, st, st1, st2)
    assert np.array_equal(p, psi) and np.array_equal(n1, psi) and np.array_equal(n2, psi1)
    for c in (0, 2):
        assert np.array_equal(s1[c], st[c]) and np.array_equal(s2[c], st1[c])
    for new, old in ((s1[1], st[1]), (s2[1], st1[1])):
        assert np.all(new[:, 0, :] == 0) and np.all(new[:, ny, :] == 0)
        assert np.array_equal(new[:, 1:ny, :], old[:, 1:ny, :])
    _, _, _, t1, t2 = O.update_pressure(g, psi, psi1, psi2, None, st1, st2)
    assert all(np.all(a == 0) for a in t1)


# ---------------------------------------------------------------- material properties
def _trilinear(go, a):
    """a0 + a1 x + a2 y + a3 z + a4 xy + a5 yz + a6 xz + a7 xyz at the cell centres."""
    n = O.shape(go)
    xs = [O.axis_coords(n[d], go.L[d], go.stretch[d], go.gamma, 0)[0] for d in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    return a[0] + a[1] * X + a[2] * Y + a[3] * Z + a[4] * X * Y + a[5] * Y * Z + a[6] * X * Z + a[7] * X * Y * Z


def test_material_props_constant_and_mixture():
    go = O.desc((12, 10, 8), stretch=(0, 1, 0), gamma=1.5)
    c = np.full(O.shape(go), 0.37)
    cs, rho, mu = O.material_props(go, c, rho_l=1.0, rho_g=0.05, mu_l=2.0, mu_g=0.1)
    assert np.allclose(cs, 0.37, rtol=0, atol=1e-15)  # the weights of both passes sum to the volume
    assert np.allclose(rho, 0.05 * 0.37 + 0.63 * 1.0, rtol=1e-15)
    assert np.allclose(mu, 0.1 * 0.37 + 0.63 * 2.0, rtol=1e-15)


def test_material_props_reproduce_trilinear_fields_inside():
    """Both interpolations are trilinear, so a trilinear field comes back
    unchanged wherever the two passes see no ghost (stretched on every axis, so
    a swapped weight or vertex shows)."""
    go = O.desc((14, 12, 10), stretch=(1, 1, 1), gamma=1.3)
    c = _trilinear(go, [0.3, 0.05, 0.1, -0.04, 0.02, -0.03, 0.01, 0.005])
    cs, _, _ = O.material_props(go, c)
    inner = (slice(2, -2),) * 3
    assert np.abs(cs[inner] - c[inner]).max() < 1e-13
    assert np.abs(cs - c).max() > 1e-4  # the ghosts (wrap, wall clamp) do change the border


def test_material_props_wall_symmetry():
    go = O.desc((8, 12, 6), stretch=(0, 1, 0), gamma=1.5)
    c = np.random.default_rng(3).random(O.shape(go))
    c = 0.5 * (c + c[:, ::-1, :])  # symmetric about the channel centre plane
    cs, _, _ = O.material_props(go, c)
    assert np.abs(cs - cs[:, ::-1, :]).max() < 1e-14
    assert np.abs(cs - c).max() > 1e-3
