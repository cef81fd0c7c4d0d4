"""tbfg_max_divergence_noalloc_dev (the continuity error without device allocation)
against tbfg_max_divergence_dev, both through the C ABI: bit-equal results on two
calls in a row, the second on a smaller grid whose maximum is smaller (so a max word
left over from the first call would show), and the axis limit of its scratch."""
import numpy as np
import pytest

import tbfsolver_amd as T
from tbfsolver_amd import _lib
from tbfsolver_amd import fields as F
from conftest import to_dev

pytestmark = pytest.mark.gpu


def velocity(torch, g, seed, scale):
    return [to_dev(a * scale, torch) for a in F.make_velocity(g, seed)]


def test_noalloc_equals_existing_entry_point(torch_gpu):
    cases = [((128, 96, 64), 1, 10.0, 11), ((32, 24, 16), 0, 1.0, 12)]
    prev = None
    for n, stretch, scale, seed in cases:
        print("case noalloc max divergence", n, stretch, flush=True)
        g = T.channel_grid(n, stretch_y=stretch, gamma=1.5)
        u = velocity(torch_gpu, g, seed, scale)
        want = F.max_divergence(g, *u)
        got = F.max_divergence_noalloc(g, *u)
        assert got == want and np.isfinite(got) and got > 0, (n, got, want)
        if prev is not None:
            assert got < prev   # the first call's larger maximum is not carried over
        prev = got
    # a non-finite face gives +inf in both
    u[0][0, 0, 0] = float("nan")
    assert F.max_divergence_noalloc(g, *u) == F.max_divergence(g, *u) == float("inf")


def test_noalloc_rejects_axes_beyond_its_scratch(torch_gpu):
    print("case noalloc limit", flush=True)
    u = [torch_gpu.zeros(8, dtype=torch_gpu.float64, device="cuda") for _ in range(3)]
    with pytest.raises(T.TbfgError) as e:
        F.max_divergence_noalloc(T.channel_grid((4098, 4, 4)), *u)   # rejected before any launch
    assert e.value.code == _lib.TBFG_EINVAL and "4096" in str(e.value)
