"""Material properties on the device (updateMaterialProps + smoothVF,
VOF.f90:197-262) against the oracle: the smoothed volume fraction, rho and mu
to round-off (1e-14 of the values' scale) on a stretched channel grid."""
import numpy as np
import pytest

import tbfsolver_amd as T
from tbfsolver_amd import fields as F
from oracle import pyoracle as O
from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,stretch", [((48, 40, 32), 1), ((32, 16, 24), 0), ((256, 192, 256), 1)])
def test_material_props_match_oracle(torch_gpu, n, stretch):
    print("case material props", n, stretch, flush=True)
    g = T.channel_grid(n, stretch_y=stretch, gamma=1.5)
    go = O.desc(n, stretch=(0, stretch, 0), gamma=1.5)
    rng = np.random.default_rng(9)
    c = np.clip(rng.random(n) * 1.4 - 0.2, 0.0, 1.0)  # with runs of pure 0 and 1
    cs_r, rho_r, mu_r = O.material_props(go, c, rho_l=1.0, rho_g=0.05, mu_l=1.0, mu_g=0.02)
    dc = to_dev(c, torch_gpu)
    cs, rho, mu = (torch_gpu.zeros_like(dc) for _ in range(3))
    F.updateMaterialProps(g, dc, cs, rho, mu, rho_l=1.0, rho_g=0.05, mu_l=1.0, mu_g=0.02)
    torch_gpu.cuda.synchronize()
    for a, b in ((cs, cs_r), (rho, rho_r), (mu, mu_r)):
        assert np.abs(to_host(a) - b).max() <= 1e-14 * max(1.0, np.abs(b).max())
    # without mu, and the second call reuses the cached coordinate tables
    rho2 = torch_gpu.zeros_like(dc)
    F.updateMaterialProps(g, dc, cs, rho2, None, rho_l=1.0, rho_g=0.05)
    torch_gpu.cuda.synchronize()
    assert torch_gpu.equal(rho2, rho)
