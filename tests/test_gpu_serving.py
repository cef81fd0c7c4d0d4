"""Serving loop on the GPU: every result of a request served in the mix
(several worker streams, requests arriving while others run) is bit for bit
the result of the same instance's requests run alone, also when the HBM
budget forces evictions and re-admissions; projected velocities are
divergence free."""
import math

import numpy as np
import pytest

from tbfsolver_amd import fields as F
from tbfsolver_amd import serving as S
from oracle import pyoracle as O
from conftest import to_host

pytestmark = pytest.mark.gpu

def pick_specs():
    """Ten instances of the serving workload's own list (serving_instances(256, 1004)):
    its first, the 256x192x256 channel, another of at least 192^3 cells, and the
    eight smallest others (both kinds), renumbered 0..9 for the request stream."""
    specs = S.serving_instances(256, 1004)
    big = [x for x in specs[1:] if x[1][0] * x[1][1] * x[1][2] >= 192 ** 3][:1]
    rest = sorted((x for x in specs[1:] if x not in big), key=lambda x: (x[1][0] * x[1][1] * x[1][2], x[0]))
    small = [x for x in rest if x[2] == "FAST"][:3] + [x for x in rest if x[2] == "MG"][:5]
    chosen = [specs[0]] + big + small
    return [(k, n, kind, nb) for k, (_, n, kind, nb) in enumerate(chosen)]


SPECS = pick_specs()


def div_bound(r):
    """Bound on max|div u| after an instance's last projection.  The grids are uniform, so
    for the PCG + MG instances (relative residual <= tolPCG = 1e-8) max|div u| <= ||div u||_2
    <= 1e-8 ||div u*||_2, with u* the velocity the last step projected (twice that here for
    round-off); the FFT instances are exact to round-off, 1e-10 max|div u*|."""
    c = math.sin(0.7 * (r.steps - 1) + 0.3 * r.index)
    ustar = [to_host(ub + c * up) for ub, up in zip(r.ub, r.up)]
    d = O.divergence(O.desc(r.n), *ustar)
    return 2e-8 * np.linalg.norm(d) if r.kind == "MG" else 1e-10 * np.abs(d).max()


def serve(workers, budget=200e9, rate=2000.0):
    reqs = S.poisson_arrivals(rate, 40 / rate, len(SPECS), seed=11)
    loop = S.ServingLoop(SPECS, precision="mixed", workers=workers, hbm_budget=budget, keep_results=True)
    out = loop.run(reqs)
    res = {k: v.cpu() for k, v in loop.results.items()}
    divs = {i: (F.max_divergence(r.grid, *r.u), div_bound(r)) for i, r in loop.res.items()}
    ev = loop.evictions
    loop.close()
    return reqs, out, res, divs, ev


def test_results_in_the_mix_equal_solo(torch_gpu):
    print("case serving mix vs solo", [x[1:3] for x in SPECS], flush=True)
    assert SPECS[0][1] == (256, 192, 256) and max(x[1][0] * x[1][1] * x[1][2] for x in SPECS[1:]) >= 192 ** 3
    reqs, out_mix, mix, divs, _ = serve(workers=4)
    assert out_mix["requests"] == len(reqs) and len(reqs) >= 20
    assert all(l is not None and l > 0 for l in out_mix["latency_ms"])
    print("case serving solo", flush=True)
    _, out_solo, solo, _, _ = serve(workers=1)
    assert set(mix) == set(solo) and len(mix) == len(reqs)
    for key in solo:
        assert torch_gpu.equal(mix[key], solo[key]), key
    steps = {}
    for _, i in reqs:
        steps[i] = steps.get(i, 0) + 1
    assert all((i, steps[i] - 1) in mix for i in steps)
    for i, (d, bound) in divs.items():
        assert np.isfinite(d) and d <= bound, (i, d, bound)


def test_evictions_keep_results(torch_gpu):
    print("case serving with evictions", flush=True)
    # room for the largest instance and a few small ones: the small ones are evicted and re-admitted
    budget = S.estimate_footprint((256, 192, 256), "MG", "mixed") + 3 * S.estimate_footprint((64, 64, 32), "MG",
                                                                                               "mixed")
    reqs, _, evicted, _, ev = serve(workers=2, budget=budget)
    assert ev > 0
    _, _, solo, _, _ = serve(workers=1)
    assert set(evicted) == set(solo)
    for key in solo:
        assert torch_gpu.equal(evicted[key], solo[key]), key
