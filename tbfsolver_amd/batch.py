"""Many independent channel instances on one GPU -- the batched workload of the
benchmark family (64 uneven instances: sizes drawn from a seed, FFT solves with
constant density on even indices, PCG + multigrid with bubbles on odd ones).

The reference runs one instance per process.  Here every instance is its own
solver handle with its own device fields; a pool of host threads drives the
solves on separate HIP streams (default priority), so the small launches of
different instances overlap on the GPU.  The library releases the Python
interpreter lock during every call, and distinct handles are independent.

  config3_instances(count, seed)  the instance list (sizes, kind, bubbles)
  Instance                        one instance: grid, inputs, handle, solve()
  run_batch(instances, workers)   all solves from a shared queue, wall time
  BatchGraph(instances, branches) every solve of a pass captured in one HIP
                                  graph: one launch per pass (replay())

One launch per pass: a persistent kernel over dependent phases would need a
grid-wide barrier between them, and on this part a device-scope barrier costs
3-20 us (it grows with the workgroups taking part) against ~2.5 us for a
dependent kernel launch.  So the pass is one graph instead: each instance's
solve is a chain of kernels in it (its PCG iterations queued on the device,
the stop test on the device), the chains spread over parallel branches.
"""
import concurrent.futures as cf
import time

import numpy as np

from . import fields as F
from .solvers import FastPoissonSolver, PCG, channel_grid

NXZ = (32, 64, 96, 128, 192)
NY = (32, 64, 96, 128)
_M64 = (1 << 64) - 1


def _uniform01(seed, idx):
    """The library's counter-based generator (splitmix64 finaliser of seed*phi + idx + 1)."""
    z = (seed * 0x9E3779B97F4A7C15 + idx + 1) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    z ^= z >> 31
    return (z >> 11) * (1.0 / 9007199254740992.0)


def config3_instances(count=64, seed=1003):
    """[(index, (nx, ny, nz), kind, bubbles)]: nx, nz from {32, 64, 96, 128, 192},
    ny from {32, 64, 96, 128}; kind "FAST" on even indices, "MG" with 4..24
    bubbles on odd ones."""
    out = []
    for i in range(count):
        u = [_uniform01(seed, 4 * i + m) for m in range(4)]
        n = (NXZ[int(u[0] * len(NXZ))], NY[int(u[1] * len(NY))], NXZ[int(u[2] * len(NXZ))])
        kind = "FAST" if i % 2 == 0 else "MG"
        nb = 4 + int(u[3] * 21) if kind == "MG" else 0
        out.append((i, n, kind, nb))
    return out


class Instance:
    """One channel instance: uniform grid (Lx = 2 pi, Ly = 2, Lz = pi), the
    seeded velocity's divergence as right-hand side, and its solver.
    precision_used records the precision the solve runs in: the requested one
    for PCG + MG instances, always "fp64" for FFT instances."""

    def __init__(self, index, n, kind, nbubbles, seed=1003, precision="mixed", tol=1e-8, adt=1e-3):
        import torch

        self.index, self.n, self.kind = index, tuple(n), kind
        self.grid = channel_grid(n)
        iseed = seed * 1000 + index
        ux, uy, uz = F.make_velocity(self.grid, iseed)
        dev = [torch.from_numpy(np.ascontiguousarray(a.T)).cuda() for a in (ux, uy, uz)]
        self.q = torch.zeros(n[2], n[1], n[0], dtype=torch.float64, device="cuda")
        F.computeSource(self.grid, dev[0], dev[1], dev[2], self.q)
        self.psi = torch.zeros_like(self.q)
        if kind == "FAST":
            # constant density: s = div u / dt made mean free, as in config 0
            self.q /= adt
            self.q -= self.q.mean()
            # FP64 whatever the precision asked: on MI355X the FP64 FFT solve is
            # 2.4-3.6x faster than the FP32 one with FP64 refinement at every
            # instance size (0.059 vs 0.214 ms at 96x64x96), and exact
            self.solver = FastPoissonSolver(self.grid, precision="fp64")
            self.precision_used = "fp64"
            self.beta = None
        else:
            rho, _ = F.make_bubble_density(self.grid, nbubbles, 0.3, iseed, nref=8)
            self.rho = torch.from_numpy(np.ascontiguousarray(rho.T)).cuda()
            self.beta = torch.zeros_like(self.q)
            F.computeBeta(self.grid, adt, self.rho, self.beta)
            self.solver = PCG(self.grid, nLevels=4, tolPCG=tol, maxIterPCG=200, precision=precision, flexible=1,
                              coarse_max_iter=0 if precision == "mixed" else -1)
            self.precision_used = precision
        self.iterations = 0
        torch.cuda.synchronize()

    def solve(self, stream=None):
        self.psi.zero_()
        if self.kind == "FAST":
            self.solver.solveFPS(self.psi, self.q, stream=stream)
        else:
            self.solver.solvePCG(self.psi, self.beta, self.q, stream=stream)
            self.iterations = self.solver.iterations
        return self.psi

    def solve_queued(self, iterations, stream=None):
        """The solve without host synchronisation (the PCG with `iterations`
        iterations queued): capturable in a graph."""
        self.psi.zero_()
        if self.kind == "FAST":
            self.solver.solveFPS(self.psi, self.q, stream=stream)
        else:
            self.solver.solvePCG_queued(self.psi, self.beta, self.q, iterations, stream=stream)
        return self.psi

    def close(self):
        self.solver.close()


def _cost(inst):
    """Relative work of a solve: PCG+MG iterations stream the grid several times
    more than the five passes of an FFT solve."""
    cells = inst.n[0] * inst.n[1] * inst.n[2]
    return cells * (1 if inst.kind == "FAST" else 8)


def run_batch(instances, workers=8):
    """Solves every instance once on `workers` host threads, each with its own
    stream; returns the wall time in seconds.  The threads take instances from
    one shared queue, largest expected work first, so no stream idles while
    another still holds several solves."""
    import threading

    import torch

    streams = [torch.cuda.Stream() for _ in range(max(1, workers))]
    order = sorted(range(len(instances)), key=lambda i: -_cost(instances[i]))
    lock = threading.Lock()
    nxt = [0]
    torch.cuda.synchronize()

    def work(w):
        s = streams[w]
        with torch.cuda.stream(s):
            while True:
                with lock:
                    if nxt[0] >= len(order):
                        break
                    i = order[nxt[0]]
                    nxt[0] += 1
                instances[i].solve(stream=s)
        s.synchronize()

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=len(streams)) as ex:
        list(ex.map(work, range(len(streams))))
    torch.cuda.synchronize()
    return time.perf_counter() - t0


class BatchGraph:
    """Every instance's solve in one HIP graph, replayed by one launch per pass.

    Each PCG + MG instance is solved once on the host's usual path first; its
    iteration count is what its chain in the graph queues (plus `margin`), and the
    device's stop test still ends the solve, so a replay gives that solve bit for
    bit.  The chains are spread over `branches` parallel branches, largest work
    first onto the least loaded branch.  harvest() reads every PCG outcome after a
    replay and returns the instances that did not converge (none while the inputs
    are the ones the counts came from)."""

    def __init__(self, instances, branches=12, margin=0):
        import torch

        self.instances = instances
        for inst in instances:
            inst.solve()
        torch.cuda.synchronize()
        self.queue = {inst.index: inst.iterations + margin for inst in instances if inst.kind == "MG"}
        load = [0] * max(1, branches)
        self.branches = [[] for _ in load]
        for i in sorted(range(len(instances)), key=lambda i: -_cost(instances[i])):
            b = load.index(min(load))
            self.branches[b].append(instances[i])
            load[b] += _cost(instances[i]) * (self.queue.get(instances[i].index, 1))
        self.graph = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream()
        side = [torch.cuda.Stream() for _ in self.branches]
        torch.cuda.synchronize()
        with torch.cuda.graph(self.graph, stream=cap):
            for chain, s in zip(self.branches, side):
                s.wait_stream(cap)
                with torch.cuda.stream(s):
                    for inst in chain:
                        inst.solve_queued(self.queue.get(inst.index, 0), stream=s)
            for s in side:
                cap.wait_stream(s)
        torch.cuda.synchronize()

    def replay(self):
        """One pass: a single graph launch (asynchronous on the current stream)."""
        self.graph.replay()

    def harvest(self):
        """After a replay has finished: every PCG instance's iterations, residual and
        convergence from its device state; returns the instances not converged."""
        bad = []
        for inst in self.instances:
            if inst.kind == "MG":
                if not inst.solver.result():
                    bad.append(inst.index)
                inst.iterations = inst.solver.iterations
        return bad

    def run(self, reps=1):
        """Wall time in seconds of `reps` passes (replays), from an idle device."""
        import torch

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def close(self):
        self.graph.reset()
