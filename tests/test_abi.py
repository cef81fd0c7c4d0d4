"""The C-ABI library loads, exports every symbol include/tbfgpu.h declares, and
its host-side logic (argument validation, input mirrors, parameter reader)
works without a GPU.  No compute kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tbfsolver_amd as T
from tbfsolver_amd import _lib
from tbfsolver_amd import fields as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "tbfgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tbfg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(L, name), name
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures out of step with the header"


def test_version_and_status():
    assert b"gfx950" in T.lib().tbfg_version()
    assert T.lib().tbfg_last_error() is not None


def test_default_params_are_the_reference_channel_values():
    p = T.default_params()
    assert (p.n_levels, p.n_pre, p.n_post, p.mg_cycles) == (4, 2, 2, 1)
    assert p.precision == 0 and p.flexible == -1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built HIP library every entry point raises: no host fallback
    stands in for the device path."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libtbfgpu.so"))
    g = T.channel_grid((64, 64, 64))
    for make in (lambda: T.FastPoissonSolver(g), lambda: T.PCG(g, nLevels=4),
                 lambda: F.max_divergence(g, None, None, None), lambda: _lib.lib()):
        with pytest.raises(RuntimeError, match="not built"):
            make()


def test_invalid_grids_and_hierarchies_are_rejected():
    g = T.channel_grid((96, 64, 100))
    with pytest.raises(T.TbfgError) as e:
        T.PCG(g, nLevels=4)          # 100 not divisible by 8 (multiGrid.f90:116-145)
    assert e.value.code == _lib.TBFG_EINVAL and "non-integer" in str(e.value)
    with pytest.raises(T.TbfgError):
        T.PCG(T.channel_grid((64, 64, 64)), nLevels=1)
    with pytest.raises(T.TbfgError):
        T.PCG(T.channel_grid((64, 64, 64)), nPreSweep=0, nPostSweep=0)
    with pytest.raises(T.TbfgError):
        T.FastPoissonSolver(T.channel_grid((100, 64, 64)))   # 100 = 4 * 25 has a factor 5
    with pytest.raises(T.TbfgError):
        T.FastPoissonSolver(T.channel_grid((63, 64, 64)))
    bad = T.channel_grid((64, 64, 64))
    bad.stretch[0] = 1                                       # periodic axes must be uniform
    with pytest.raises(T.TbfgError):
        T.FastPoissonSolver(bad)


def test_velocity_mirror():
    g = T.channel_grid((16, 12, 8))
    ux, uy, uz = F.make_velocity(g, 1000)
    ux2, uy2, uz2 = F.make_velocity(g, 1000)
    assert np.array_equal(ux, ux2) and np.array_equal(uy, uy2)
    assert np.all(uy[:, 0, :] == 0) and np.all(uy[:, -1, :] == 0)
    allv = np.concatenate([ux.ravel(), uz.ravel(), uy[:, 1:-1, :].ravel()])
    assert abs(allv.mean()) < 0.05 and abs(allv.std() - 1) < 0.05
    ux3, _, _ = F.make_velocity(g, 1001)
    assert not np.array_equal(ux, ux3)


@pytest.mark.parametrize("stretch", [0, 1])
def test_bubble_density_mirror(stretch):
    g = T.channel_grid((64, 64, 64), stretch_y=stretch, gamma=1.5)
    rho, cen = F.make_bubble_density(g, 18, 0.3, 1001, nref=20, rho_l=1.0, rho_g=0.05)
    assert rho.min() >= 0.05 - 1e-12 and rho.max() <= 1.0 + 1e-12
    # non-overlapping (minimum image in x, z), centres at least R from the walls
    R = 0.15
    assert np.all(cen[:, 1] >= R) and np.all(cen[:, 1] <= 2.0 - R)
    for a in range(18):
        for b in range(a):
            d = np.abs(cen[a] - cen[b])
            d[0] = min(d[0], 2 * np.pi - d[0])
            d[2] = min(d[2], np.pi - d[2])
            assert np.sum(d ** 2) >= 4 * R * R - 1e-12
    # gas volume = 18 spheres (the smoothing conserves it on uniform grids)
    xf = [np.linspace(0, L, n + 1) for L, n in zip((2 * np.pi, 2.0, np.pi), (64, 64, 64))]
    if stretch:
        j = np.arange(65)
        xf[1] = 0.5 * 2.0 * (1 + np.tanh(-1.5 + 2 * 1.5 * j / 64) / np.tanh(1.5))
    V = np.diff(xf[0])[:, None, None] * np.diff(xf[1])[None, :, None] * np.diff(xf[2])[None, None, :]
    cs = (1.0 - rho) / (1.0 - 0.05)
    gas = np.sum(V * cs)
    exact = 18 * 4.0 / 3.0 * np.pi * R ** 3
    assert abs(gas - exact) / exact < (0.03 if stretch else 0.02)


def test_parameter_file_reader(tmp_path):
    f = tmp_path / "pcg_solver"
    f.write_text("! comment\nnLevels 5\nnPreSweep 2\nnPostSweep 4\ntolMG 4.d-1\nmaxIterMG 100\n"
                 "fullInfoMG .F.\ntolPCG 1.d-10\nmaxIterPCG 100\n")
    prm = T.read_parameter_file(str(f))
    assert prm == {"nLevels": 5, "nPreSweep": 2, "nPostSweep": 4, "tolMG": 0.4, "maxIterMG": 100,
                   "fullInfoMG": False, "tolPCG": 1e-10, "maxIterPCG": 100}


def test_fft_pass_division_by_multiply_high():
    """The FFT passes divide butterfly indices (< 2^15) by the pass's
    sub-lengths with mulhi(x, ceil(2^32 / d)) instead of an integer division:
    exact for every divisor that a length the solver accepts (2^a 3^b, even,
    <= 4096) produces -- n / radix and the partial radix products."""
    def radices(n):
        out, m = [], n
        for r in (4, 2, 3):
            while m % r == 0:
                out.append(r)
                m //= r
        return out if m == 1 else None

    divisors = set()
    for n in range(2, 4097, 2):
        rad = radices(n)
        if rad is None:
            continue
        ns = 1
        for r in rad:
            divisors.update((n // r, ns))
            ns *= r
    x = np.arange(1 << 15, dtype=np.uint64)
    for d in sorted(divisors):
        if d == 1:
            continue  # the passes use the identity for d = 1
        magic = np.uint64(((1 << 32) + d - 1) // d)
        assert np.array_equal((x * magic) >> np.uint64(32), x // np.uint64(d)), d


def test_precision_codes_match_header_and_are_validated():
    """The precision codes of the header and of the ctypes binding agree; the
    Python mirror rejects names it does not know, the C ABI codes it does not
    know (before any device call)."""
    txt = open(os.path.join(ROOT, "include", "tbfgpu.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define (TBFG_PREC_[A-Z0-9_]+) (\d+)", txt)}
    assert defs == {"TBFG_PREC_FP64": _lib.PREC_FP64, "TBFG_PREC_MIXED": _lib.PREC_MIXED,
                    "TBFG_PREC_MIXED_BF16": _lib.PREC_MIXED_BF16}
    g = T.channel_grid((64, 64, 64))
    for bad in ("mixed16", "bf16", "FP64"):
        with pytest.raises(ValueError):
            T.PCG(g, precision=bad)
    with pytest.raises(ValueError):
        T.FastPoissonSolver(g, precision="mixed-bf16")   # the FFT solver has no bf16 form
    p = T.default_params()
    p.precision = 7
    h = C.c_void_p()
    assert T.lib().tbfg_pcg_create(C.byref(g), C.byref(p), C.byref(h)) == _lib.TBFG_EINVAL
    assert b"precision" in T.lib().tbfg_last_error()
