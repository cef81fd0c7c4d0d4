"""Serving: an open-loop stream of projection-step requests over channel
instances kept resident on the GPU, with continuous batching.

Each request is one projection step of one instance (the right-hand side from
the instance's predicted velocity u*, the pressure solve warm-started from its
last pressure, the projection, the old-level rotation: poissonEqn.f90:347-459
and momentumEqn.f90:795-1127 through PoissonEqn).  The reference runs one
instance per process and has no request loop; this module is the engine's own.

  serving_instances(count, seed)     instance list: sizes up to 256x192x256, 70 % PCG + MG, 30 % FFT
  poisson_arrivals(rate, duration)   open-loop arrival times (exponential gaps, seeded)
  HbmAdmission(budget)               residency accounting by device bytes, LRU eviction
  Resident                           one instance on the device: fields, solver handle, step()
  ServingLoop(specs, ...).run(reqs)  the request loop: latency per request, throughput

Continuous batching: `workers` host threads, each with its own HIP stream
(default priority), take the earliest arrived request whose instance is idle,
so requests arriving while others run are started as soon as a slot frees and
their kernels overlap on the GPU.  Requests of one instance run in arrival
order, one at a time, so an instance's results are those of running its
requests alone.  Every kernel reduces in a fixed order, so a result is the
same bit for bit alone and in the mix.

Admission by HBM footprint: an instance is made resident on its first request
if its device bytes fit the budget next to those already resident; otherwise
idle instances are evicted, least recently used first (their pressure state
is kept in host memory and restored on re-admission), and the request waits
while the resident instances that are busy hold the budget.
"""
import collections
import math
import threading
import time

import numpy as np

from . import fields as F
from .batch import NXZ, NY, _uniform01
from .poisson import PoissonEqn
from .solvers import channel_grid

ALPHA_DT = 1e-3


# sizes of the served instances: the batched workload's sets plus the largest
# channel, up to 256 x 192 x 256
SERVE_NXZ = NXZ + (256,)
SERVE_NY = NY + (192,)


def serving_instances(count=256, seed=1004, p_mg=0.7):
    """[(index, (nx, ny, nz), kind, bubbles)]: nx, nz from {32, 64, 96, 128, 192, 256},
    ny from {32, 64, 96, 128, 192}, except instance 0, which is the largest channel
    (256 x 192 x 256) so that every list spans the whole size range; kind "MG"
    (4..24 bubbles) with probability p_mg, else "FAST"."""
    out = []
    for i in range(count):
        u = [_uniform01(seed, 5 * i + m) for m in range(5)]
        n = (SERVE_NXZ[int(u[0] * len(SERVE_NXZ))], SERVE_NY[int(u[1] * len(SERVE_NY))],
             SERVE_NXZ[int(u[2] * len(SERVE_NXZ))])
        if i == 0:
            n = (SERVE_NXZ[-1], SERVE_NY[-1], SERVE_NXZ[-1])
        kind = "MG" if u[3] < p_mg else "FAST"
        nb = 4 + int(u[4] * 21) if kind == "MG" else 0
        out.append((i, n, kind, nb))
    return out


def poisson_arrivals(rate, duration, n_instances, seed=7):
    """[(t, instance)] with exponential gaps of mean 1/rate over [0, duration),
    instances drawn uniformly."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append((t, int(rng.integers(n_instances))))


def estimate_footprint(n, kind, precision):
    """Device bytes of a Resident before it exists (admission decides on this;
    the accounting then uses the measured value)."""
    cells = n[0] * n[1] * n[2]
    fields = 8 * (5 + 9 + 3) * cells  # psi, p, rho, s, beta; u*, u base, u perturbation; old levels
    per_cell = {("MG", "mixed"): 56, ("MG", "fp64"): 80, ("FAST", "fp64"): 56}
    return fields + per_cell[(kind, "fp64" if kind == "FAST" else precision)] * cells + (8 << 20)


class HbmAdmission:
    """Residency accounting: bytes held per resident instance, least recently
    used order, and the pick of idle instances to evict for a newcomer."""

    def __init__(self, budget):
        self.budget = int(budget)
        self.held = collections.OrderedDict()  # instance -> bytes, least recently used first

    @property
    def used(self):
        return sum(self.held.values())

    def touch(self, idx):
        self.held.move_to_end(idx)

    def victims(self, need, busy):
        """Idle residents to evict (LRU first) so that `need` more bytes fit, or
        None if even evicting every idle resident would not make room."""
        if need > self.budget:
            raise ValueError("an instance of %d bytes exceeds the HBM budget of %d bytes" % (need, self.budget))
        free = self.budget - self.used
        out = []
        for idx, b in self.held.items():
            if free >= need:
                break
            if idx in busy:
                continue
            out.append(idx)
            free += b
        return out if free >= need else None

    def add(self, idx, nbytes):
        self.held[idx] = int(nbytes)

    def remove(self, idx):
        return self.held.pop(idx)


class Resident:
    """One channel instance on the device (uniform grid, Lx = 2 pi, Ly = 2,
    Lz = pi): density, base and perturbation velocity, pressure and its old
    levels, the pressure-equation object with its solver handle.  The
    velocities are drawn on the device (N(0, 1) faces from a generator seeded
    by the instance, v = 0 on the walls); the bubble density on the host."""

    def __init__(self, spec, precision="mixed", seed=1004, state=None):
        import torch

        index, n, kind, nb = spec
        self.index, self.n, self.kind = index, tuple(n), kind
        self.grid = channel_grid(self.n)
        iseed = seed * 1000 + index

        def dev(a):
            return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda()

        gen = torch.Generator(device="cuda")
        gen.manual_seed(iseed)
        nx, ny, nz = self.n

        def velocity(scale):
            u = [torch.randn(shape, generator=gen, dtype=torch.float64, device="cuda") * scale
                 for shape in ((nz, ny, nx), (nz, ny + 1, nx), (nz, ny, nx))]
            u[1][:, 0, :] = 0.0
            u[1][:, ny, :] = 0.0
            return u

        self.ub = velocity(1.0)
        self.up = velocity(0.1)
        self.u = [torch.empty_like(a) for a in self.ub]
        if kind == "MG":
            rho, _ = F.make_bubble_density(self.grid, nb, 0.3, iseed, nref=8)
            self.rho = dev(rho)
            # the same V-cycle in both precisions (its coarsest level relaxed, not solved to 1e-10)
            self.eqn = PoissonEqn(self.grid, "MG", 1.0, 1.0 / 20.0, precision=precision, nLevels=4, tolPCG=1e-8,
                                  maxIterPCG=200, flexible=1, coarse_max_iter=0)
        else:
            self.rho = torch.ones(self.n[2], self.n[1], self.n[0], dtype=torch.float64, device="cuda")
            # the FFT solve in FP64 in both modes: faster on MI355X than its FP32 form
            # with FP64 refinement, and exact
            self.eqn = PoissonEqn(self.grid, "FAST", 1.0, 1.0, precision="fp64")
        self.psi = torch.zeros_like(self.rho)
        self.p = torch.zeros_like(self.rho)
        self.steps = 0
        if state is not None:
            self._restore(state)
        torch.cuda.current_stream().synchronize()
        self.footprint = self._footprint()

    def _footprint(self):
        t = [self.rho, self.psi, self.p] + self.ub + self.up + self.u
        nbytes = sum(a.numel() * a.element_size() for a in t)
        # allocated by the first steps: s and the two old pressure levels, and beta (MG)
        nbytes += (4 if self.kind == "MG" else 3) * self.psi.numel() * 8
        solver = self.eqn.fftSolver_ if self.kind == "FAST" else self.eqn.pcgs_
        return nbytes + solver.footprint

    def step(self, stream=None):
        """One projection step: u* = u_base + c_m u_pert (c_m fixed by the step
        number), the pressure solve from the last pressure, the projection, the
        pressure update.  Returns the pressure (device)."""
        c = math.sin(0.7 * self.steps + 0.3 * self.index)
        for u, ub, up in zip(self.u, self.ub, self.up):
            u.copy_(ub)
            u.add_(up, alpha=c)
        self.eqn.solvePoissonEqn(self.psi, self.rho, *self.u, ALPHA_DT, stream=stream)
        self.eqn.makeVelocityDivFree(self.psi, self.rho, *self.u, ALPHA_DT, stream=stream)
        self.eqn.updatePressure(self.psi, self.p, stream=stream)
        self.steps += 1
        return self.psi

    def state(self):
        """Host copy of what the next step depends on (for eviction)."""
        old = [None if a is None else a.cpu() for a in getattr(self.eqn, "old_", [None, None])]
        return {"psi": self.psi.cpu(), "p": self.p.cpu(), "old": old, "steps": self.steps}

    def _restore(self, st):
        self.psi.copy_(st["psi"])
        self.p.copy_(st["p"])
        self.eqn.old_ = [None if a is None else a.cuda() for a in st["old"]]
        self.steps = st["steps"]

    def close(self):
        self.eqn.close()


class ServingLoop:
    """The request loop over `specs` (see serving_instances).  `workers`
    threads with a stream each; `hbm_budget` bytes of resident instances."""

    resident = Resident  # the instance type (the host-logic tests substitute a stand-in)

    def __init__(self, specs, precision="mixed", workers=8, hbm_budget=200e9, seed=1004, keep_results=False):
        self.specs = {s[0]: s for s in specs}
        self.precision = precision
        self.workers = max(1, int(workers))
        self.seed = seed
        self.adm = HbmAdmission(hbm_budget)
        self.res = {}        # instance -> Resident
        self.saved = {}      # instance -> host state of an evicted instance
        self.keep = keep_results
        self.results = {}    # (instance, step) -> pressure after the step (device copy), with keep_results
        self.evictions = 0
        self._adm_lock = threading.Lock()

    def preload(self, indices=None):
        """Makes instances resident ahead of the stream (in index order, while
        the budget allows); returns the number resident."""
        for i in sorted(self.specs) if indices is None else indices:
            spec = self.specs[i]
            if self.adm.used + estimate_footprint(spec[1], spec[2], self.precision) > self.adm.budget:
                break
            self._admit(i, busy=set())
        return len(self.res)

    def _admit(self, idx, busy):
        """Resident instance idx (caller holds _adm_lock), or None if the budget
        is held by busy instances."""
        if idx in self.res:
            self.adm.touch(idx)
            return self.res[idx]
        spec = self.specs[idx]
        need = estimate_footprint(spec[1], spec[2], self.precision)
        ev = self.adm.victims(need, busy)
        if ev is None:
            return None
        for v in ev:
            r = self.res.pop(v)
            self.saved[v] = r.state()
            r.close()
            self.adm.remove(v)
            self.evictions += 1
        r = self.resident(spec, self.precision, self.seed, state=self.saved.pop(idx, None))
        self.res[idx] = r
        self.adm.add(idx, r.footprint)
        return r

    def run(self, requests):
        """Serves [(t_arrival, instance)] open loop (t in seconds from the start);
        returns a dict with the latency of every request (ms, in request order),
        p50/p99, throughput and the wall time."""
        reqs = sorted(enumerate(requests), key=lambda x: x[1][0])
        lat = [None] * len(requests)
        pending = collections.deque()
        busy = set()
        nxt = [0]
        cv = threading.Condition()
        streams = self._streams()
        errors = []
        self._sync()
        t0 = time.perf_counter()

        def take():
            # (caller holds cv) the earliest arrived request whose instance is idle
            now = time.perf_counter() - t0
            while nxt[0] < len(reqs) and reqs[nxt[0]][1][0] <= now:
                pending.append(reqs[nxt[0]])
                nxt[0] += 1
            for q in pending:
                if q[1][1] not in busy:
                    pending.remove(q)
                    busy.add(q[1][1])
                    return q
            return None

        def work(w):
            s = streams[w]
            with self._on(s):
                while True:
                    with cv:
                        while True:
                            if errors:
                                return
                            q = take()
                            if q is not None:
                                break
                            if nxt[0] >= len(reqs) and not pending:
                                return
                            wait = None
                            if nxt[0] < len(reqs):
                                wait = max(0.0, reqs[nxt[0]][1][0] - (time.perf_counter() - t0))
                            cv.wait(timeout=wait if wait is not None else 0.05)
                    k, (t_arr, idx) = q
                    try:
                        while True:
                            with cv:
                                held = set(busy) - {idx}
                            with self._adm_lock:
                                r = self._admit(idx, held)
                            if r is not None:
                                break
                            with cv:  # the budget is held by busy instances: wait for one to finish
                                cv.wait(timeout=0.01)
                        step = r.steps
                        psi = r.step(stream=s)
                        if self.keep:
                            self.results[(idx, step)] = psi.clone()
                        self._wait(s)
                    except Exception as e:  # surfaced after the loop
                        with cv:
                            errors.append(e)
                            busy.discard(idx)
                            cv.notify_all()
                        return
                    done = time.perf_counter() - t0
                    with cv:
                        lat[k] = (done - t_arr) * 1e3
                        busy.discard(idx)
                        cv.notify_all()

        threads = [threading.Thread(target=work, args=(w,)) for w in range(self.workers)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        self._sync()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        arr = np.array(lat, dtype=float)
        first = min(r[0] for r in requests) if requests else 0.0
        last_done = max(r[0] + l / 1e3 for r, l in zip(requests, lat)) if requests else 0.0
        return {
            "requests": len(requests),
            "p50_ms": float(np.percentile(arr, 50)) if len(arr) else None,
            "p99_ms": float(np.percentile(arr, 99)) if len(arr) else None,
            "mean_ms": float(arr.mean()) if len(arr) else None,
            "throughput_rps": len(requests) / (last_done - first) if len(requests) > 1 else None,
            "wall_s": wall,
            "latency_ms": lat,
        }

    # the device side of the loop: one HIP stream per worker (default priority)
    def _streams(self):
        import torch

        return [torch.cuda.Stream() for _ in range(self.workers)]

    def _on(self, s):
        import torch

        return torch.cuda.stream(s)

    def _wait(self, s):
        s.synchronize()

    def _sync(self):
        import torch

        torch.cuda.synchronize()

    def close(self):
        for r in self.res.values():
            r.close()
        self.res.clear()
