"""The instance-parallel oracle pool (worker processes, used by the batched-workload test
and the benchmark's all-core CPU baseline) reproduces the in-process oracle twin exactly
(CPU only: the payloads are built from the oracle's own topology, no GPU)."""
import numpy as np

from oracle import oracle as orc
from oracle_pool import _twin, parallel_substeps, run_substeps
from xrtailor_amd import scenes
from xrtailor_amd._ffi import XrtParams


def _params(**kw):
    p = XrtParams()
    defaults = dict(damping=0.98, relaxation_factor=0.25, long_range_stretchiness=1.0,
                    bvh_tolerance=0.003, frame_time=1.0 / 60.0, num_substeps=10,
                    num_iterations=4, num_collision_passes=10, solve_bending=1,
                    enable_ground=1, ground_height=-1.0, pbd_thickness=0.002,
                    repulsion_relaxation=0.25, pair_update_interval=10, bend_model=1,
                    enable_obstacle_contact=1, enable_self_contact=1,
                    pair_capacity_per_face=64)
    defaults.update(kw)
    p.gravity[1] = -9.81
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def _payload(spec, xp, seed):
    edges = orc.build_edges(spec.tris)
    bends = orc.build_bends(spec.tris, edges)
    scol, sn = orc.color(edges, len(spec.pos), seed, 0)
    bcol, bn = orc.color(bends, len(spec.pos), seed, 0)
    v = np.zeros_like(spec.x0, np.float64)
    v[:, 1] = -0.5
    return dict(x0=spec.x0.astype(np.float64), v=v, w=np.ones(len(spec.pos)),
                edges=edges, rest=orc.rest_lengths(spec.pos, edges).astype(np.float32).astype(np.float64),
                s_comp=np.full(len(edges), spec.stretch_compliance), scol=scol, ns=sn,
                bends=bends, theta=orc.rest_angles(spec.pos, bends).astype(np.float32).astype(np.float64),
                b_comp=np.full(len(bends), spec.bend_compliance), bcol=bcol, nb=bn, lra=None,
                pos=spec.pos, tris=spec.tris, plain_edges=edges,
                obstacle=(np.asarray(spec.obstacle[0], np.float32), spec.obstacle[1]),
                obstacle_x=None, params=bytes(xp))


def test_parallel_pool_equals_in_process_twins():
    xp = _params()
    specs = scenes.batched(seed=9, n_instances=3, side_lo=8, side_hi=14)
    for sp in specs:                                  # obstacle top just below the cloth
        sp.obstacle[0][:, 1] += -0.002 - sp.obstacle[0][:, 1].max()
    payloads = [_payload(sp, xp, 3) for sp in specs]
    out, busy = parallel_substeps(payloads, n_substeps=2, workers=2)
    assert busy > 0.0
    for p, (x, sec, pid) in zip(payloads, out):
        ref, _, _ = run_substeps(p, 2)
        np.testing.assert_array_equal(x, ref)
        cl, sc, _, _ = _twin(p)
        assert sc is not None and sc.has_obs
    moved = max(float(np.abs(x - p["x0"]).max()) for p, (x, _, _) in zip(payloads, out))
    assert moved > 1e-4
