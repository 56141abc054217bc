"""FP64 oracle twins of solver instances run in worker processes (test infrastructure and
the benchmark's CPU baseline only).

instance_payload() gathers, in the process that owns the GPU solver, everything an oracle
twin of one instance needs (topology, colors, LRA CSR, FP32-rounded start state, obstacle)
as plain arrays; run_substeps() rebuilds the twin in a worker and advances it.  Workers
are started with `spawn` and import numpy and the oracle only: they never touch the GPU."""
from __future__ import annotations

import os
import time
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np


def instance_payload(solver, inst, spec, params, x0=None, v=None, obstacle=True,
                     obstacle_x=None) -> dict:
    """Oracle-twin inputs of instance `inst` (start state: x0 / v, default the spec's x0 at
    rest; obstacle_x: the obstacle's frame-end positions, default its rest positions)."""
    n = solver.counts(inst)
    edges, rest, scol = solver.edges(inst)
    bends, theta, bcol = solver.bends(inst)
    x0 = spec.x0 if x0 is None else x0
    lra = solver.lra(inst) if n.lra_slots else None
    obs = None
    if obstacle and spec.obstacle is not None:
        obs = (np.asarray(spec.obstacle[0], np.float32), np.asarray(spec.obstacle[1], np.int32))
    return dict(
        x0=np.asarray(x0, np.float64), v=np.zeros_like(x0, np.float64) if v is None else
        np.asarray(v, np.float64), w=solver.inv_mass(inst).astype(np.float64),
        edges=edges, rest=rest.astype(np.float64),
        s_comp=np.broadcast_to(np.asarray(spec.stretch_compliance, np.float64), (n.n_edges,)).copy(),
        scol=scol, ns=n.stretch_colors, bends=bends, theta=theta.astype(np.float64),
        b_comp=np.full(n.n_bends, float(spec.bend_compliance)), bcol=bcol, nb=n.bend_colors,
        lra=lra, pos=np.asarray(spec.pos), tris=np.asarray(spec.tris, np.int32),
        plain_edges=solver.edges(inst, colors=False)[0], obstacle=obs,
        obstacle_x=None if obstacle_x is None else np.asarray(obstacle_x, np.float32),
        params=bytes(params))


def _twin(p):
    from oracle import oracle as orc
    from oracle import scene as osc
    from xrtailor_amd._ffi import XrtParams

    xp = XrtParams.from_buffer_copy(p["params"])
    kw = {}
    if p["lra"] is not None:
        sv, ptr, sl, rs = p["lra"]
        kw = dict(slot_vert=sv, lra_ptr=ptr, lra_slot=sl, lra_rest=rs.astype(np.float64))
    cl = orc.Cloth(p["x0"], p["v"], p["w"], p["edges"], p["rest"], p["s_comp"], p["scol"], p["ns"],
                   p["bends"], p["theta"], p["b_comp"], p["bcol"], p["nb"], **kw)
    op, dt = orc.params_from(xp)
    contacts = p["obstacle"] is not None or xp.enable_self_contact
    if not contacts:
        return cl, None, op, dt
    sc = osc.ContactScene(cl, p["pos"], p["tris"], p["plain_edges"], op, xp, obstacle=p["obstacle"])
    if p["obstacle_x"] is not None:
        sc.o_x = p["obstacle_x"].copy()
    return cl, sc, op, dt


def run_substeps(p: dict, n_substeps: int = 1):
    """(x0 after n_substeps as float64 (nv, 3), seconds spent in the substeps, worker pid)."""
    cl, sc, op, dt = _twin(p)
    t0 = time.perf_counter()
    for _ in range(n_substeps):
        if sc is None:
            cl.substep(op, dt)
        else:
            sc.substep()
    sec = time.perf_counter() - t0
    return cl.state()[0].copy(), sec, os.getpid()


def default_workers() -> int:
    """Worker processes: the host's cores, at most 16 (one GPU's share of a GPU host)."""
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                      else (os.cpu_count() or 1)))


def parallel_substeps(payloads, n_substeps: int = 1, workers: int | None = None):
    """run_substeps over every payload, instance-parallel in `workers` processes.
    Returns (list of (x, seconds, pid) in payload order, critical path): the critical path
    is the largest per-worker sum of substep seconds, i.e. the time the instance-parallel
    substeps take when each worker's twins are set up beforehand (set-up is not timed)."""
    workers = workers or default_workers()
    ctx = mp.get_context("spawn")
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        list(ex.map(_warm, range(workers)))           # workers started and imported
        out = list(ex.map(run_substeps, payloads, [n_substeps] * len(payloads)))
    busy: dict = {}
    for _, sec, pid in out:
        busy[pid] = busy.get(pid, 0.0) + sec
    return out, max(busy.values()) if busy else 0.0


def _warm(_):
    from oracle import scene  # noqa: F401
    return os.getpid()
