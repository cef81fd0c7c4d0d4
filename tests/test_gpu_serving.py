"""Serving loop on the GPU: every result of a request served in the mix
(several worker streams, requests arriving while others run) is bit for bit
the result of the same instance's requests run alone, also when the HBM
budget forces evictions and re-admissions; projected velocities are
divergence free."""
import numpy as np
import pytest

from tbfsolver_amd import fields as F
from tbfsolver_amd import serving as S

pytestmark = pytest.mark.gpu

SPECS = [
    (0, (32, 32, 32), "MG", 6),
    (1, (64, 32, 32), "FAST", 0),
    (2, (32, 64, 64), "MG", 9),
    (3, (64, 32, 64), "MG", 4),
    (4, (32, 32, 64), "FAST", 0),
    (5, (64, 64, 32), "MG", 12),
    (6, (96, 32, 32), "FAST", 0),
    (7, (32, 96, 32), "MG", 5),
]


def serve(workers, budget=200e9, rate=4000.0):
    reqs = S.poisson_arrivals(rate, 40 / rate, len(SPECS), seed=11)
    loop = S.ServingLoop(SPECS, precision="mixed", workers=workers, hbm_budget=budget, keep_results=True)
    out = loop.run(reqs)
    res = {k: v.cpu() for k, v in loop.results.items()}
    divs = {i: F.max_divergence(r.grid, *r.u) for i, r in loop.res.items()}
    ev = loop.evictions
    loop.close()
    return reqs, out, res, divs, ev


def test_results_in_the_mix_equal_solo(torch_gpu):
    print("case serving mix vs solo", flush=True)
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
    for i, d in divs.items():
        assert np.isfinite(d) and d < 1e-5, (i, d)  # |div u| after projection (u ~ 1, h ~ 0.1)


def test_evictions_keep_results(torch_gpu):
    print("case serving with evictions", flush=True)
    budget = 3 * S.estimate_footprint((64, 64, 32), "MG", "mixed")
    reqs, _, evicted, _, ev = serve(workers=2, budget=budget)
    assert ev > 0
    _, _, solo, _, _ = serve(workers=1)
    assert set(evicted) == set(solo)
    for key in solo:
        assert torch_gpu.equal(evicted[key], solo[key]), key
