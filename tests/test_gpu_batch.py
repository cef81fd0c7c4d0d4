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
def test_config3_all_instances_concurrent_equal_solo_and_oracle(torch_gpu):
    """All 64 instances of the batched workload (every size it draws, both kinds),
    solved together on 8 streams: each equals its solo solve bit for bit (same
    iteration count), and each matches the oracle -- the FFT instances (FP64) to
    1e-11 (grids with nz = 192 and ny = 64 reach 1.4e-12: the y solve of the smallest
    wavenumbers is conditioned ~ 1/|lambda_min|), the PCG + MG instances (mixed precision, here to 1e-11) to 1e-9
    (relative L2, mean free) against the oracle's FP64 solve to 1e-12."""
    specs = Bt.config3_instances(64, 1003)
    sizes = sorted({x[1] for x in specs})
    print("case batch all 64", len(sizes), "sizes, largest", max(sizes, key=lambda n: n[0] * n[1] * n[2]), flush=True)
    insts = [Bt.Instance(*x, tol=1e-11) for x in specs]
    assert {i.precision_used for i in insts if i.kind == "FAST"} == {"fp64"}
    Bt.run_batch(insts, workers=8)
    together = [(i.iterations, to_host(i.psi)) for i in insts]
    for inst, (it, psi) in zip(insts, together):
        inst.solve()
        torch_gpu.cuda.synchronize()
        assert inst.iterations == it, inst.index
        assert np.array_equal(to_host(inst.psi), psi), inst.index
    worst = {"FAST": 0.0, "MG": 0.0}
    for inst in insts:
        go = O.desc(inst.n)
        q = to_host(inst.q)
        a = to_host(inst.psi)
        if inst.kind == "FAST":
            ref = O.fft_solve(go, q)
            e = np.linalg.norm(a - ref) / np.linalg.norm(ref)
            assert e < 1e-11, (inst.index, inst.n, e)
        else:
            b = to_host(inst.beta)
            ref, _, _, rc = O.pcg_solve(go, O.params(n_levels=4, tol=1e-12, max_iter=300, flexible=1), b, q)
            assert rc == 0
            a, ref = a - a.mean(), ref - ref.mean()
            e = np.linalg.norm(a - ref) / np.linalg.norm(ref)
            assert e < 1e-9, (inst.index, inst.n, e)
        worst[inst.kind] = max(worst[inst.kind], e)
    print("  worst rel L2 error vs oracle:", worst, flush=True)
    for inst in insts:
        inst.close()


@pytest.mark.gpu
def test_config3_one_graph_launch_per_pass_equals_solo(torch_gpu):
    """All 64 instances of the batched workload captured in one HIP graph (one
    launch per pass, 8 parallel branches; each PCG's iterations queued on the
    device, its stop test on the device).  Three replays, with eager work that
    overwrites every output and synchronisations between them: each replay equals
    the instances solved one at a time on the host-driven path, bit for bit, with
    the same iteration counts."""
    specs = Bt.config3_instances(64, 1003)
    print("case batch graph, 64 instances", flush=True)
    insts = [Bt.Instance(*x) for x in specs]
    solo = []
    for inst in insts:
        inst.solve()
        torch_gpu.cuda.synchronize()
        solo.append((inst.iterations, to_host(inst.psi)))
    bg = Bt.BatchGraph(insts, branches=8)
    assert len(bg.queue) == 32 and all(k > 0 for k in bg.queue.values())
    for rep in range(3):
        for inst in insts:
            inst.psi.fill_(float("nan"))
        torch_gpu.cuda.synchronize()
        bg.replay()
        torch_gpu.cuda.synchronize()
        assert bg.harvest() == [], rep
        for inst, (it, psi) in zip(insts, solo):
            if inst.kind == "MG":
                assert inst.iterations == it, (rep, inst.index)
            assert np.array_equal(to_host(inst.psi), psi), (rep, inst.index)
        print("  replay", rep, "equal", flush=True)
    bg.close()
    for inst in insts:
        inst.close()
