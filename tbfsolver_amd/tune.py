"""Tile-shape tuner for the multigrid kernels.

The library launches its sweep and operator kernels with a tile shape chosen
per level: from the tuning table (tbfg_tune_set) when it has an entry for the
level's cells and precision, else by a built-in order of preference.  This
module fills the table by measurement: for the finest level of a PCG+MG
instance it times whole solves with every candidate shape of one kernel
family at a time (sweep from a non-zero start, first sweep / sweep with the
dot products, FP64 operator product) and keeps the fastest.  Shapes change
only how cells are grouped into workgroups: the sweep results agree to
round-off (a cell recomputed in a tile's ring may round differently from the
same cell inside a tile), the operator's dot product differs in its summation
order only.

  candidates(kind, dims, precision)     shapes with a kernel that tile the level
  tune_instance(dims, precision, reps)  measure one instance size, enter the winners
  autotune(sizes, precision, reps)      every distinct size; returns the table
  save(table, path) / load(path)        JSON; load enters the entries

  python -m tbfsolver_amd.tune --config3 --out table.json   (the 64-instance workload's sizes)
"""
import argparse
import ctypes as C
import json
import time

import numpy as np

from ._lib import check, lib

KINDS = {"sweep": 0, "sweep_edge": 1, "op": 2}
SWEEP_F = [(64, 8, 256), (32, 16, 256), (64, 16, 512), (32, 8, 128)]
SWEEP_D = [(64, 4, 128), (64, 8, 256), (32, 8, 128)]
OP = [(128, 512, 0), (64, 512, 0), (32, 256, 0), (64, 256, 0), (-1, 0, 0)]  # (-1, 0, 0): column segments
PLANE_MARCH_MIN_CELLS = 1 << 20  # smallest level smoothed by the plane-march sweep


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def elem_bytes(precision):
    return 4 if precision == "mixed" else 8


def set_entry(kind, dims, precision, shape):
    check(lib().tbfg_tune_set(KINDS[kind], _i3(dims), elem_bytes(precision), _i3(shape)))


def get_entry(kind, dims, precision):
    out = (C.c_int32 * 3)()
    check(lib().tbfg_tune_get(KINDS[kind], _i3(dims), elem_bytes(precision), out))
    return tuple(out)


def clear():
    check(lib().tbfg_tune_clear())


def candidates(kind, dims, precision):
    nx, ny, nz = dims
    if kind == "op":
        return [s for s in OP if s[0] < 0 or (nx % s[0] == 0 and ny % (s[1] // s[0]) == 0)]
    if nz < 16 or nz % 2 or ny % 2 or nx * ny * nz < PLANE_MARCH_MIN_CELLS:
        return []  # the level is not smoothed by the plane-march sweep
    table = SWEEP_F if precision == "mixed" else SWEEP_D
    return [s for s in table if nx % s[0] == 0 and ny % s[1] == 0]


def _case(dims, seed=4321, nbubbles=12):
    import torch

    from . import fields as F
    from .solvers import channel_grid

    g = channel_grid(dims, stretch_y=True, gamma=1.5)
    rho, _ = F.make_bubble_density(g, nbubbles, 0.3, seed, nref=8)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda()

    ux, uy, uz = (dev(a) for a in F.make_velocity(g, seed))
    drho = dev(rho)
    q = torch.zeros_like(drho)
    F.computeSource(g, ux, uy, uz, q)
    beta = torch.zeros_like(drho)
    F.computeBeta(g, 1e-3, drho, beta)
    return g, q, beta


def _time_solves(g, q, beta, precision, reps):
    import torch

    from .solvers import PCG

    s = PCG(g, nLevels=4, tolPCG=1e-8, maxIterPCG=200, precision=precision, flexible=1, coarse_max_iter=0)
    psi = torch.zeros_like(q)
    s.solvePCG(psi, beta, q)  # warm-up
    times = []
    for _ in range(reps):
        psi.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.solvePCG(psi, beta, q)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    s.close()
    return float(np.median(times)) * 1e3


def tune_instance(dims, precision="mixed", reps=3, log=None):
    """Times each candidate of each family on a PCG+MG instance of these cells
    (the other families at their current choice) and enters the fastest.
    Returns {kind: {"shape": [...], "ms": t, "tried": {shape: ms}}}."""
    dims = tuple(int(v) for v in dims)
    g, q, beta = _case(dims)
    out = {}
    for kind in ("op", "sweep", "sweep_edge"):
        cands = candidates(kind, dims, precision)
        if len(cands) < 2:
            continue
        tried = {}
        for shape in cands:
            set_entry(kind, dims, precision, shape)
            tried[shape] = _time_solves(g, q, beta, precision, reps)
        best = min(tried, key=tried.get)
        set_entry(kind, dims, precision, best)
        out[kind] = {"shape": list(best), "ms": round(tried[best], 4),
                     "tried": {"x".join(str(v) for v in k): round(t, 4) for k, t in tried.items()}}
        if log:
            log("tune %s %s %s: %s (%.3f ms; %s)" % ("x".join(map(str, dims)), precision, kind, best, tried[best],
                                                    ", ".join("%s %.3f" % (k, t) for k, t in tried.items())))
    return out


def autotune(sizes, precision="mixed", reps=3, log=None):
    """The table for every distinct size: {"nx,ny,nz": {kind: {...}}}."""
    table = {"precision": precision, "entries": {}}
    for dims in sorted({tuple(int(v) for v in d) for d in sizes}):
        table["entries"][",".join(map(str, dims))] = tune_instance(dims, precision, reps, log)
    return table


def save(table, path):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path, apply=True):
    """Reads a table written by save() and enters its winners."""
    with open(path) as f:
        table = json.load(f)
    if apply:
        prec = table.get("precision", "mixed")
        for key, kinds in table["entries"].items():
            dims = tuple(int(v) for v in key.split(","))
            for kind, e in kinds.items():
                set_entry(kind, dims, prec, e["shape"])
    return table


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--config3", action="store_true", help="the PCG+MG sizes of the 64-instance workload")
    ap.add_argument("--size", action="append", default=[], help="nx,ny,nz (repeatable)")
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    sizes = [tuple(int(v) for v in s.split(",")) for s in args.size]
    if args.config3:
        from .batch import config3_instances

        sizes += [n for _, n, kind, _ in config3_instances(64, 1003) if kind == "MG"]
    table = autotune(sizes, args.precision, args.reps, log=lambda m: print(m, flush=True))
    save(table, args.out)
    print("wrote", args.out, len(table["entries"]), "sizes", flush=True)


if __name__ == "__main__":
    main()
