"""Batched uneven instances (the 64-instance workload): the instance list is
deterministic, and solves driven concurrently on several streams equal the
same instances solved one at a time, and the oracle."""
import numpy as np
import pytest

from tbfsolver_amd import batch as Bt
from oracle import pyoracle as O
from conftest import to_host


def test_config3_instance_list():
    inst = Bt.config3_instances(64, 1003)
    assert len(inst) == 64
    assert inst == Bt.config3_instances(64, 1003)
    for i, n, kind, nb in inst:
        assert n[0] in Bt.NXZ and n[2] in Bt.NXZ and n[1] in Bt.NY
        assert kind == ("FAST" if i % 2 == 0 else "MG")
        assert (4 <= nb <= 24) if kind == "MG" else nb == 0
    sizes = {n for _, n, _, _ in inst}
    assert len(sizes) > 20  # the draws spread over the size set
    assert Bt._uniform01(7, 0) != Bt._uniform01(7, 1) and 0 <= Bt._uniform01(7, 0) < 1


@pytest.mark.gpu
def test_concurrent_instances_match_solo_solves(torch_gpu):
    picks = [x for x in Bt.config3_instances(64, 1003) if x[1][0] * x[1][1] * x[1][2] <= 96 * 96 * 96][:8]
    print("case batch", [x[1] for x in picks], flush=True)
    insts = [Bt.Instance(*x) for x in picks]
    Bt.run_batch(insts, workers=4)
    together = [(i.iterations, to_host(i.psi)) for i in insts]
    for inst, (it, psi) in zip(insts, together):
        inst.solve()
        torch_gpu.cuda.synchronize()
        assert inst.iterations == it
        assert np.array_equal(to_host(inst.psi), psi)
    # two of them against the oracle
    for inst in insts[:2]:
        go = O.desc(inst.n)
        q = to_host(inst.q)
        if inst.kind == "FAST":
            ref = O.fft_solve(go, q)
            assert np.linalg.norm(to_host(inst.psi) - ref) / np.linalg.norm(ref) < 1e-10
        else:
            b = to_host(inst.beta)
            ref, _, _, rc = O.pcg_solve(go, O.params(n_levels=4, tol=1e-12, max_iter=200, flexible=1), b, q)
            a = to_host(inst.psi)
            a, ref = a - a.mean(), ref - ref.mean()
            assert np.linalg.norm(a - ref) / np.linalg.norm(ref) < 1e-6
    for inst in insts:
        inst.close()
