"""Scenes for the contact tests and the oracle twin built from oracle topology."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from oracle import scene as osc
from xrtailor_amd import scenes


def sphere_drop(n=24, side=1.0, r=0.3, height=0.305, v_down=-1.0, seed=3):
    """Cloth just above an icosphere (s3) with a downward velocity: contacts after predict."""
    pos, tris = scenes.grid(n, side, height)
    x0 = (pos + 1e-4 * scenes.Rng(seed).normal(pos.size).reshape(-1, 3)).astype(np.float32)
    spec = scenes.ClothSpec(pos, tris, np.zeros(0, np.int32), 1e-7, 1e-3, x0,
                            scenes.icosphere(3, r))
    v0 = np.zeros_like(x0)
    v0[:, 1] = v_down
    return spec, v0


def two_layers(n=16, side=0.5, gap=0.0015, wrinkle=8e-4, angle=0.45, seed=5):
    """Two cloth sheets in one mesh, gap below the repulsion thickness; the sheets are
    wrinkled (vertical jitter) so that their face boxes overlap (pairs are searched at
    thickness 1e-6), and the upper one is turned by `angle` about y so that edge pairs
    cross instead of running parallel (the edge-edge weights of near-parallel edges are
    ill-conditioned in FP32)."""
    pos, tris = scenes.grid(n, side, 0.0)
    pos2 = pos.copy()
    c, s = np.cos(angle), np.sin(angle)
    pos2[:, 0], pos2[:, 2] = c * pos[:, 0] - s * pos[:, 2], s * pos[:, 0] + c * pos[:, 2]
    pos2[:, 1] += gap
    P = np.concatenate([pos, pos2])
    T = np.concatenate([tris, tris + len(pos)]).astype(np.int32)
    jit = scenes.Rng(seed).normal(P.size).reshape(-1, 3)
    x0 = P + 1e-5 * jit
    x0[:, 1] += wrinkle * np.clip(jit[:, 1], -1.5, 1.5)
    return scenes.ClothSpec(P, T, np.zeros(0, np.int32), 1e-7, 1e-3, x0.astype(np.float32))


def oracle_scene(spec, xparams, v0=None, seed=0, solver=None, inst=0):
    """Oracle twin.  With a solver: its topology, colors and FP32 state; else the oracle's own."""
    op, _ = orc.params_from(xparams)
    if solver is not None:
        from gpu_helpers import oracle_of
        cl = oracle_of(solver, inst, spec.stretch_compliance, spec.bend_compliance)
        edges = solver.edges(inst, colors=False)[0]
    else:
        edges = orc.build_edges(spec.tris)
        bends = orc.build_bends(spec.tris, edges)
        rest = orc.rest_lengths(spec.pos, edges)
        theta = orc.rest_angles(spec.pos, bends)
        scol, sn = orc.color(edges, len(spec.pos), seed, xparams.coloring)
        bcol, bn = orc.color(bends, len(spec.pos), seed, xparams.coloring)
        w = np.ones(len(spec.pos))
        w[spec.pinned] = 0.0
        v = np.zeros_like(spec.x0) if v0 is None else v0
        cl = orc.Cloth(spec.x0.astype(np.float64), np.asarray(v, np.float64), w, edges,
                       rest.astype(np.float32).astype(np.float64),
                       np.full(len(edges), spec.stretch_compliance), scol, sn, bends,
                       theta.astype(np.float32).astype(np.float64),
                       np.full(len(bends), spec.bend_compliance), bcol, bn)
    return osc.ContactScene(cl, spec.pos, spec.tris, edges, op, xparams,
                            obstacle=None if spec.obstacle is None else
                            (np.asarray(spec.obstacle[0], np.float32), spec.obstacle[1]))
