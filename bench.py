#!/usr/bin/env python3
"""Benchmark of the cloth substep on one MI355X.

Workload (default): 256 independent cloth instances, grid side round(32*6^u) (log-uniform
in [32, 192]), per-instance stretch / bend compliance log-uniform in [1e-9, 1e-6] /
[1e-5, 1e-3], 10 substeps x 10 iterations per frame at 60 fps, ground plane, FP32,
synthetic from seed.  A step is one frame (10 substeps) of all instances.

value        batched vertex-substeps/s (one hipGraph per frame; every contact phase and
             constraint color of all instances per launch; without contacts the same wide
             launches or one persistent launch per substep, whichever the autotuner times
             faster), inputs resident in HBM
plain_value  the same workload one instance at a time through per-instance launches
roofline     dominant kernel, per-launch device time from HIP events on the solver stream:
             with contacts the EE repulsion item kernel, algorithmic bytes per launch
             = 8 (EE items) + 16 N + 8 E (items, positions, edges read once); without
             contacts the persistent substep kernel (or the wide substep's launches
             together), sum_i (64 N_i + 16 E_i + 24 B_i + 12 K_i); HBM peak 8 TB/s.  traffic: HBM bytes per launch from the committed
             rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that kernel (profiles/traffic.json),
             null when absent
cpu_baseline the serial FP64 oracle on one core, one substep per instance over a bounded
             sample; the same leg reports the single-substep relative error of the GPU's
             batched substep against it (parity at the measured configuration)

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
                    [--workload batched|swift|quality|garment|garment_body|hanging]
                    (swift: the same 256 cloths with body contact and the particle
                     self-collision of the reference's swift mode instead of PBD repulsion;
                     roofline kernel ParticleCollideKernel, algorithmic bytes per launch
                     48 N + 4 (neighbor entries + N))
                    (quality: the same 256 cloths with body and self contact and the
                     quality mode's FEM constraints (Green strain over faces with moduli
                     1e3, 1e3, 4e2 and Poisson ratios 0.3, isometric bending stiffness 1e-2
                     over hinges) in place of the XPBD stretch and bend; roofline kernel
                     WideStrainKernel, algorithmic bytes
                     per launch 128 per face of the color (32 B element, three float4
                     vertices read and written), the isometric bending launches beside it
                     at 172 B per hinge; with --no-contacts the persistent kernel, 64 N +
                     32 F + 44 hinges + 12 K per launch)
                    (garment: 64 tube garments on capsule-chain bodies skinned on the GPU
                     every frame (LBS from joint transforms), bindings at the top ring and
                     waist, LRA, body contact; a step is the skinning plus one frame, timed
                     by the host clock; roofline kernel LbsKernel, algorithmic bytes 36 per
                     body vertex + 8 per weight entry + 48 per joint)
                    (garment_body: one 160 x 125 tube garment (20,000 vertices, the shoulder
                     ring's 160 vertices as LRA slots) on a 6,890-vertex capsule-chain body
                     skinned by 24 joints every frame, body and self contact; roofline kernel
                     the EE repulsion kernel, as for the batched workload)
                    [--no-graph]   (un-captured substeps, for profiler runs)
                    [--rollout-frames 120] [--plain-at 0,40,80]
                                   (batched workload: the whole 120-frame rollout, per-frame
                                    times, with the plain path run at the listed frames)
                    [--no-tune]    (default launch shapes: no autotune before timing)
    python bench.py --copy-calibration   (5 device copies of 256 MiB: the known-size kernel
                                          that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def substep_kernel_name(solver) -> str:
    """The batched substep without contacts is one persistent launch, or (when the
    autotuner chose it) the wide substep's launches, timed together."""
    if solver.tuning()["path"] == 0:
        return "BatchedSubstepKernel"
    return "wide substep (all launches of one substep)"


def measured_traffic(kernel: str, workload: str, algorithmic: float):
    """HBM bytes per launch of `kernel` on `workload` from the committed PMC summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy calibration;
    the counters need their own profiler runs, so they cannot come from this run).  Only a
    record whose algorithmic bytes per launch match this run's within 2 % is used (the
    record describes the same launches); otherwise null."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rec.get("kernels", []):
        if r.get("kernel") == kernel and r.get("workload") == workload:
            alg = r.get("algorithmic_bytes_per_launch")
            if not alg or abs(alg - algorithmic) > 0.02 * algorithmic:
                return None
            return r.get("bytes_per_launch")
    return None


def algorithmic_bytes(solver, insts) -> tuple[int, int]:
    """(bytes per substep launch, vertices): state read+write (64 B/vertex), stretch
    16 B, bend 24 B and LRA 12 B per constraint."""
    total, nv = 0, 0
    for i in insts:
        c = solver.counts(i)
        total += 64 * c.nv + 16 * c.n_edges + 24 * c.n_bends + 12 * c.lra_constraints
        nv += c.nv
    return total, nv


def substep_bytes(solver, insts, specs, params, pairs: int, contacts: bool) -> dict:
    """Compulsory HBM bytes of one substep of every instance (the whole-substep roofline):
    state read + written once (64 B per vertex), constraint records read once (16 B per
    stretch, 24 B per bend, 12 B per LRA constraint) and, with contacts, the obstacle
    positions (16 B per obstacle vertex), the LBVH builds (40 B per node, 2F - 1 nodes; the
    frame-start rebuilds amortized over the frame's substeps), the obstacle refits of the
    collision passes and the cloth-tree updates of the pair refreshes, 8 B per face pair
    (written and read back) and the contact records of each collision pass (16 B per VF-DCD
    and VF-CCD record, i.e. per cloth vertex, 48 B per EE record, i.e. per cloth edge).
    Intermediates (the iterate, multipliers, Jacobi sums, repulsion test items) are not
    counted."""
    n = [solver.counts(i) for i in insts]
    N = sum(c.nv for c in n)
    E = sum(c.n_edges for c in n)
    B = sum(c.n_bends for c in n)
    K = sum(c.lra_constraints for c in n)
    F = sum(c.nt for c in n)
    dyn = 64 * N + 16 * E + 24 * B + 12 * K
    out = {"vertices": N, "edges": E, "bends": B, "lra": K, "faces": F, "dynamics": dyn,
           "contact": 0}
    if contacts:
        obs = [sp.obstacle for sp in specs if sp.obstacle is not None]
        n_obs = sum(len(o[0]) for o in obs)
        f_obs = sum(len(o[1]) for o in obs)
        obs_tree = 40 * (2 * f_obs - len(obs))              # every obstacle tree's nodes
        cloth_tree = 40 * (2 * F - len(n))                  # every cloth tree's nodes
        ns, ni = params.num_substeps, params.num_iterations
        passes = params.num_collision_passes / ns          # collision passes per substep
        self_c = bool(params.enable_self_contact)
        refreshes = (ni + params.pair_update_interval - 1) // params.pair_update_interval
        trees = obs_tree * (1.0 / ns + passes)              # frame build + midstep refits
        if self_c:
            trees += cloth_tree * (1.0 / ns + refreshes)    # frame build + pair-refresh updates
        records = passes * (2 * 16 * N + 48 * E)
        contact = 16 * n_obs + trees + records + (8 * pairs * refreshes if self_c else 0)
        out.update({"obstacle_vertices": n_obs, "obstacle_faces": f_obs, "pairs": pairs,
                    "contact": contact})
    out["bytes_per_substep"] = out["dynamics"] + out["contact"]
    out["bytes_per_vertex_substep"] = out["bytes_per_substep"] / N
    return out


def rollout_leg(solver, insts, specs, frames: int, plain_at) -> dict:
    """The workload's whole rollout from its initial state: `frames` frames, each timed with
    HIP events on the solver stream.  The frames in `plain_at` run the plain path instead of
    the captured batched frame; both paths give identical states, so the trajectory is the
    same either way and the plain samples spread over the rollout."""
    import numpy as np

    for i, sp in zip(insts, specs):
        solver.SetState(i, sp.x0, np.zeros_like(sp.x0))
    bms, pms = [], []
    for f in range(frames):
        if f in plain_at:
            pms.append(solver.time_frames(False, 1))
        else:
            bms.append(solver.time_frames(True, 1))
        if (f + 1) % 20 == 0:
            print(f"[bench] rollout: {f + 1} / {frames} frames", file=sys.stderr, flush=True)
    vs = sum(len(sp.pos) for sp in specs) * solver.params.num_substeps
    b = np.array(bms)
    return {"frames": frames, "batched_frames": len(bms), "plain_frames_sampled": sorted(plain_at),
            "batched_ms_per_frame": {"mean": float(b.mean()), "min": float(b.min()),
                                     "median": float(np.median(b)), "max": float(b.max())},
            "plain_ms_per_frame": [float(x) for x in pms],
            "batched_value": vs * len(bms) / (b.sum() * 1e-3),
            "plain_value": vs * len(pms) / (sum(pms) * 1e-3) if pms else None,
            "batched_over_plain": float(np.mean(pms) / b.mean()) if pms else None}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_and_parity(specs, params, seed, contacts, budget_s=12.0, before_start=None,
                            twin_setup=None, parallel=True):
    """CPU baseline leg and parity.  The GPU's batched substep from the initial state is
    compared with the FP64 oracle twin of each instance (relative error).
    * serial: the oracle (1 core) advances instances one after another until the time
      budget is spent (the reported cpu_baseline);
    * parallel (no twin set-up hook): every instance, instance-parallel over the host's
      cores (one worker process per core, at most 16), timed by the busiest worker's
      substep seconds (set-up excluded), so parity covers every instance."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from gpu_helpers import oracle_of, rel_err
    from oracle import oracle as orc
    from oracle import scene as osc
    from oracle_pool import default_workers, instance_payload, parallel_substeps
    from xrtailor_amd.solver import build_solver

    solver, insts = build_solver(specs, params, seed=seed, obstacles=contacts,
                                 before_start=before_start)
    solver.SubstepBatched()
    solver.sync()
    gpu = [solver.GetState(i)[0] for i in insts]
    payloads = None
    if parallel and twin_setup is None:
        payloads = [instance_payload(solver, i, sp, params, obstacle=contacts)
                    for i, sp in zip(insts, specs)]
    op, dt = orc.params_from(params)
    done_vs, t_used, n_inst, worst = 0, 0.0, 0, 0.0
    t_wall = time.perf_counter()
    for i, sp in zip(insts, specs):
        oc = oracle_of(solver, i, sp.stretch_compliance, sp.bend_compliance,
                       x0=sp.x0, v=np.zeros_like(sp.x0))
        if contacts or twin_setup is not None:
            edges = solver.edges(i, colors=False)[0]
            obs = None if sp.obstacle is None or not contacts else (
                np.asarray(sp.obstacle[0], np.float32), sp.obstacle[1])
            twin = osc.ContactScene(oc, sp.pos, sp.tris, edges, op, params, obstacle=obs)
            if twin_setup is not None:
                twin_setup(n_inst, twin)
            t0 = time.perf_counter()
            twin.substep()
        else:
            t0 = time.perf_counter()
            oc.substep(op, dt)
        t_used += time.perf_counter() - t0
        worst = max(worst, rel_err(gpu[n_inst], oc.state()[0]))
        done_vs += oc.nv
        n_inst += 1
        if n_inst % 16 == 0:
            print(f"[bench] cpu leg: {n_inst} instances, {t_used:.1f} s timed", file=sys.stderr,
                  flush=True)
        # the twins' set-up is not timed: the leg also ends after 4 x the budget of wall time
        if t_used >= budget_s or time.perf_counter() - t_wall >= 4 * budget_s:
            break
    solver.close()
    what = "without contacts"
    if contacts:
        what = ("with body contact and particle self-collision" if params.enable_particle_collision
                else "with body and self contact")
    if params.constraint_model == 1:
        what += ", FEM constraints"
    model, ncpu = cpu_model(), os.cpu_count()
    base = {"value": done_vs / t_used, "unit": "vertex-substeps/s", "cores": 1, "kind": "port",
            "sample": f"one substep ({what}) of the first {n_inst} of {len(specs)} instances "
                      f"({done_vs} vertex-substeps, {t_used:.1f} s), serial FP64 oracle; "
                      f"host CPU {model}, {ncpu} cores present"}
    par = None
    if payloads is not None:
        workers = default_workers()
        print(f"[bench] cpu leg: every instance on {workers} worker processes", file=sys.stderr,
              flush=True)
        out, busy = parallel_substeps(payloads, 1, workers)
        errs = [rel_err(g, x) for g, (x, _, _) in zip(gpu, out)]
        worst, n_inst = max(errs), len(errs)
        vs = sum(len(sp.pos) for sp in specs)
        par = {"value": vs / busy, "unit": "vertex-substeps/s", "cores": workers, "kind": "port",
               "sample": f"one substep ({what}) of all {len(specs)} instances ({vs} vertex-"
                         f"substeps), FP64 oracle instance-parallel on {workers} worker "
                         f"processes, timed by the busiest worker ({busy:.1f} s of substeps, "
                         f"set-up excluded); host CPU {model}, {ncpu} cores present"}
    return base, par, worst, n_inst


def adjoint_leg(frames: int = 60, checkpoint_every: int = 10, n: int = 128, seed: int = 5) -> dict:
    """Differentiable rollout: n x n cloth dropped on a sphere (body contact on), `frames`
    frames with a checkpoint every `checkpoint_every`.  Loss = mean squared distance of the
    final vertices to a target made by a forward run with per-edge compliances perturbed
    log-normally (sigma 0.5).  Reports the per-substep cost of the adjoint (taped forward +
    reverse) against the forward substep on the same wide launches, and the whole
    rollout's backward (checkpoint recompute included) against its forward."""
    import numpy as np

    from xrtailor_amd import scenes
    from xrtailor_amd.solver import build_solver, make_params

    spec = scenes.differentiable_drop(seed=seed, n=n)
    params = make_params(bend_model=1, enable_obstacle_contact=1, ground_height=-1.0)
    solver, (i,) = build_solver([spec], params, seed=seed)
    ne = solver.counts(i).n_edges
    z = scenes.Rng(seed + 1).normal(ne)
    base = np.full(ne, spec.stretch_compliance)
    solver.SetStretchCompliance(i, (base * np.exp(0.5 * z)).astype(np.float32))
    solver.AdjointBegin(frames, frames)
    target, _ = solver.GetState(i)
    solver.SetStretchCompliance(i, base.astype(np.float32))
    solver.SetState(i, spec.x0, np.zeros_like(spec.x0))
    ms_fwd, ms_tape, ms_rev = solver.time_adjoint(20)
    t0 = time.perf_counter()
    solver.AdjointBegin(frames, checkpoint_every)
    t_fwd = time.perf_counter() - t0
    x, _ = solver.GetState(i)
    nv = len(x)
    loss = float(((x.astype(np.float64) - target) ** 2).sum(axis=1).mean())
    solver.AdjointSeed(i, (2.0 / nv) * (x - target), None)
    t0 = time.perf_counter()
    solver.AdjointBackward()
    t_bwd = time.perf_counter() - t0
    gs, _ = solver.GradCompliance(i)
    _, gv = solver.GradState(i)
    solver.close()
    return {"workload": f"{n}x{n} cloth over a sphere (icosphere s4 r 0.3), body contact, "
                        f"{frames} frames, checkpoint every {checkpoint_every}",
            "substep_ms": {"forward": ms_fwd, "taped_forward": ms_tape, "reverse": ms_rev},
            "substep_ratio": (ms_tape + ms_rev) / ms_fwd,
            "rollout_forward_s": t_fwd, "rollout_backward_s": t_bwd,
            "rollout_ratio": t_bwd / t_fwd, "loss": loss,
            "grad_finite": bool(np.isfinite(gs).all() and np.isfinite(gv).all())}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "swift", "quality", "garment", "garment_body", "hanging"])
    ap.add_argument("--garments", type=int, default=64, help="instances of --workload garment")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--pair-capacity", type=int, default=256,
                    help="face pairs per cloth face of the self-contact pair lists")
    ap.add_argument("--vf-items", type=int, default=128,
                    help="VF repulsion test items per cloth face of the batched item lists")
    ap.add_argument("--ee-items", type=int, default=256,
                    help="EE repulsion test items per cloth face of the batched item lists")
    ap.add_argument("--bend-model", default="consistent", choices=["consistent", "reference"],
                    help="bend projection: the consistent model (default) or the reference's "
                         "formula verbatim (same cost; it gains energy at small compliance)")
    ap.add_argument("--bend-compliance", type=float, default=None,
                    help="one bend compliance for every instance (the reference ships 100)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--plain-frames", type=int, default=1)
    ap.add_argument("--rollout-frames", type=int, default=120,
                    help="batched workload: time the whole rollout (0: skip)")
    ap.add_argument("--plain-at", default="0,40,80",
                    help="rollout frames run by the plain path instead (comma separated)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (no autotune)")
    ap.add_argument("--no-cpu-parallel", action="store_true",
                    help="skip the instance-parallel oracle leg (parity over the serial sample)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain path (profiling runs)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable rollout")
    ap.add_argument("--no-contacts", action="store_true",
                    help="batched workload without its obstacles and self contact")
    ap.add_argument("--no-graph", action="store_true",
                    help="time un-captured batched substeps instead of graph replays")
    ap.add_argument("--record", action="store_true",
                    help="also time the same frames with the per-frame cache (positions and "
                         "normals copied out every frame, copy overlapped with compute)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only 5 device copies of 256 MiB (PMC byte-counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = 64 << 20
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": 4 * n, "copies": 5}), flush=True)
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU only (instances are independent; "
                                   "N GPUs would each run their own share): use --gpus 1"}))
        return 2

    import numpy as np

    from xrtailor_amd import scenes
    from xrtailor_amd.solver import build_solver, make_params

    gs, before = None, None
    if a.workload == "garment":
        gs = scenes.garments(seed=a.seed, n_instances=a.garments)
        specs = [g[0] for g in gs]
        wname = (f"garment {a.garments} tubes (ring 64-128 x rows 40-80) on capsule-chain bodies "
                 f"(10,242 vertices, 12 joints) skinned on the GPU each frame, top ring and waist "
                 f"bound to the body, LRA, body contact, seed {a.seed}")

        def before(s, insts):
            for i, g in zip(insts, gs):
                s.SetObstacleSkinning(i, *g[1].skinning(), g[1].amp.shape[0])
                s.AddBindings(i, g[2], g[3], g[4])
    elif a.workload == "garment_body":
        gs = [scenes.garment_on_body(seed=a.seed)]
        specs = [gs[0][0]]
        wname = ("garment on an animated body: 160 x 125 tube (20,000 vertices) with its shoulder "
                 "ring as 160 LRA slots and bound to a 6,890-vertex capsule-chain body skinned by "
                 f"24 joints on the GPU each frame, body and self contact, seed {a.seed}")

        def before(s, insts):
            for i, g in zip(insts, gs):
                s.SetObstacleSkinning(i, *g[1].skinning(), g[1].amp.shape[0])
                s.AddBindings(i, g[2], g[3], g[4])
    elif a.workload in ("batched", "swift", "quality"):
        specs = scenes.batched(seed=a.seed, n_instances=a.instances)
        wname = f"batched {a.instances} cloths, side round(32*6^u) in [32,192], seed {a.seed}"
    else:
        specs = [scenes.hanging_cloth(seed=a.seed)]
        wname = "hanging cloth 64x64, 1 m, two corners pinned"
    swift = a.workload == "swift"
    quality = a.workload == "quality"
    contacts = (a.workload in ("batched", "swift", "quality", "garment", "garment_body")
                and not a.no_contacts)
    if a.bend_compliance is not None:
        for sp in specs:
            sp.bend_compliance = a.bend_compliance
        wname += f", bend compliance {a.bend_compliance:g}"
    garment = a.workload in ("garment", "garment_body")
    bound = garment
    self_contact = contacts and not swift and a.workload != "garment"
    # quality: moduli large enough that the strain projection acts from small strains on
    # (it is skipped while sum_i w_i |grad_i|^2 <= 1e-6, which at the reference's default
    # moduli of 1 holds up to strains of 4-20 % on these meshes), bending stiffness of the
    # reference's quality example configuration
    fem = dict(constraint_model=1, fem_xx_stiffness=1e3, fem_yy_stiffness=1e3,
               fem_xy_stiffness=4e2, fem_bending_stiffness=1e-2) if quality else {}
    # face-pair lists: the 256 cloths draped over their obstacles reach ~76 pairs per face
    # within the first frames and 130 in the crumpled frames near the end of the 120-frame
    # rollout; 256 per face holds every pair of the whole rollout (an overflow would make
    # the timing calls fail with XRT_ERR_CAPACITY).  Test-item lists of twice the default
    # capacities keep the most crumpled cloth on the item path (with 64 / 128 one instance
    # evaluated per pair); results are the same either way.
    bend_model = 1 if a.bend_model == "consistent" else 0
    params = make_params(bend_model=bend_model, ground_height=-1.0,
                         enable_obstacle_contact=int(contacts),
                         enable_self_contact=int(self_contact),
                         enable_particle_collision=int(swift),
                         pair_capacity_per_face=a.pair_capacity,
                         repulsion_vf_items_per_face=a.vf_items,
                         repulsion_ee_items_per_face=a.ee_items, **fem)
    if quality:
        wname += (", sphere / capsule obstacle and self contact" if contacts else "") + \
            ", FEM strain (moduli 1e3/1e3/4e2) and isometric bending (quality mode)"
    elif swift:
        wname += (", sphere / capsule obstacle" if contacts else "") + \
            ", particle self-collision (swift mode)"
    elif contacts and a.workload == "batched":
        wname += ", sphere / capsule obstacle and self contact"
    print(f"[bench] building {len(specs)} instances", file=sys.stderr, flush=True)
    solver, insts = build_solver(specs, params, seed=a.seed, obstacles=contacts,
                                 before_start=before)
    bytes_sub, n_verts = algorithmic_bytes(solver, insts)
    # launch shapes: the project's autotuner times its candidates on these instances (state
    # restored afterwards) before anything is timed
    tuning = None if a.no_tune else solver.autotune(3)
    if tuning:
        print(f"[bench] autotune: {tuning}", file=sys.stderr, flush=True)
    # garment: a step skins every body (joint transforms of frame f, uploaded) then runs the
    # frame; the transforms of every frame are prepared before the timed region
    T = None
    if garment:
        nf = a.warmup + a.steps + 2 + a.plain_frames
        T = [np.concatenate([g[1].transforms((f + 1) / 60.0).reshape(-1) for g in gs])
             for f in range(nf)]
    info = solver.device_info()

    # batched: warmup, then K timed frames (graph replays), HIP events on the solver stream
    t0 = time.perf_counter()
    if garment:
        def frame(f):
            solver.SkinObstacles([T[f]])
            if a.no_graph:
                for _ in range(params.num_substeps):
                    solver.SubstepBatched()
            else:
                solver.FramesBatched(1)
        for f in range(a.warmup):
            frame(f)
        solver.sync()
        t0 = time.perf_counter()
        for f in range(a.warmup, a.warmup + a.steps):
            frame(f)
        solver.sync()
        ms = (time.perf_counter() - t0) * 1e3
    elif a.no_graph:
        for _ in range(a.warmup * params.num_substeps):
            solver.SubstepBatched()
        solver.sync()
        t0 = time.perf_counter()
        nsub = a.steps * params.num_substeps
        ms = solver.time_batched_substep(nsub) * nsub
    else:
        solver.FramesBatched(a.warmup)
        solver.sync()
        t0 = time.perf_counter()
        ms = solver.time_frames(True, a.steps)
    wall = time.perf_counter() - t0
    value = n_verts * params.num_substeps * a.steps / (ms * 1e-3)
    print(f"[bench] batched {ms:.2f} ms / {a.steps} frames (wall {wall:.3f} s)",
          file=sys.stderr, flush=True)

    record = None
    if a.record and not garment:
        # the per-frame cache: positions + normals of every vertex, 24 B per vertex per frame
        # bracketed by unrecorded frames just before and after (the cost per frame changes
        # as the simulation advances)
        solver.RecordFrames(1, normals=True)
        before = solver.time_frames(True, a.steps)
        t1 = time.perf_counter()
        solver.RecordFrames(a.steps, normals=True)
        rms = (time.perf_counter() - t1) * 1e3
        after = solver.time_frames(True, a.steps)
        base = 0.5 * (before + after)
        record = {"ms_per_frame": rms / a.steps, "frame_ms_before": before / a.steps,
                  "frame_ms_after": after / a.steps, "overhead": rms / base - 1.0,
                  "bytes_per_frame": 24 * n_verts,
                  "d2h_gbs": 24 * n_verts * a.steps / (rms * 1e-3) / 1e9}
        print(f"[bench] recorded frames {rms / a.steps:.1f} ms/frame vs {before / a.steps:.1f} "
              f"before and {after / a.steps:.1f} after", file=sys.stderr, flush=True)

    # dominant kernel: per-launch device time from HIP events around each launch
    rep_counts, particle, binding, fem = None, None, None, None
    if quality:
        kernel = "WideStrainKernel"
        strain_ms, iso_ms = solver.time_fem(2 * params.num_iterations)
        faces = sum(solver.counts(i).nt for i in insts)
        hinges = sum(solver.counts(i).n_bends for i in insts)
        f_launch = max(solver.fem(i)["n_colors"] for i in insts)
        b_launch = max(solver.counts(i).bend_colors for i in insts)
        kern_ms = strain_ms
        # one face color of every instance per launch: element (32 B) read, three vertices
        # read and written (96 B), averaged over the colors
        bytes_kern = 128 * faces / f_launch
        # hinge: indices (16 B), weights K (16 B), coeff (4 B), multiplier read and written
        # (8 B), four vertices read and written (128 B)
        fem = {"strain_ms": strain_ms, "strain_launches_per_pass": f_launch, "faces": faces,
               "iso_bend_ms": iso_ms, "iso_bend_launches_per_pass": b_launch, "hinges": hinges,
               "iso_bend_gbs": 172 * hinges / b_launch / (iso_ms * 1e-3) / 1e9}
        if not contacts:
            # the frame is the persistent kernel: state read + written (64 B per vertex),
            # strain elements 32 B, isometric data 44 B per hinge, LRA 12 B, read once
            kernel = substep_kernel_name(solver)
            kern_ms = solver.time_batched_substep(max(10, a.steps))
            lra = sum(solver.counts(i).lra_constraints for i in insts)
            bytes_kern = 64 * n_verts + 32 * faces + 44 * hinges + 12 * lra
    elif a.workload == "garment":
        kernel = "LbsKernel"
        kern_ms, bind_ms = solver.time_binding(20)
        bodies = sum(len(g[1].pos) for g in gs)
        entries = sum(len(g[1].skinning()[1]) for g in gs)
        joints = sum(g[1].amp.shape[0] for g in gs)
        # rest point read, skinned point written, CSR offset (36 B per body vertex), joint and
        # weight (8 B per entry), the 3x4 transform (48 B per joint)
        bytes_kern = 36 * bodies + 8 * entries + 48 * joints
        nb = sum(len(g[2]) for g in gs)
        binding = {"lbs_ms": kern_ms, "body_vertices": bodies, "weight_entries": entries,
                   "joints": joints, "binding_solve_ms": bind_ms, "bindings": nb,
                   # item (48 B), cloth vertex read + written (32 B), three body points (48 B)
                   "binding_solve_gbs": 128 * nb / (bind_ms * 1e-3) / 1e9}
    elif swift:
        kernel = "ParticleCollideKernel"
        kern_ms, hash_ms, pc = solver.time_particles(2 * params.num_substeps)
        # x, x0 read, delta written (48 B per particle), list entries and sentinels (4 B)
        bytes_kern = 48 * n_verts + 4 * (pc["entries"] + n_verts)
        particle = {"collide_ms": kern_ms, "hash_ms": hash_ms, **pc,
                    "max_num_neighbors": params.max_num_neighbors,
                    "inter_leaved_hash": params.inter_leaved_hash}
    elif contacts:
        kernel = "MultiRepItemKernel<1>"
        kern_ms, rep_counts = solver.time_repulsion(params.num_substeps)  # a whole frame
        n_edges = sum(solver.counts(i).n_edges for i in insts)
        bytes_kern = 8 * rep_counts["ee_items"] + 16 * n_verts + 8 * n_edges
    else:
        kernel = substep_kernel_name(solver)
        kern_ms = solver.time_batched_substep(max(10, a.steps))
        bytes_kern = bytes_sub
    achieved = bytes_kern / (kern_ms * 1e-3) / 1e9
    traffic = measured_traffic(kernel, wname, bytes_kern)
    sb = substep_bytes(solver, insts, specs, params,
                       rep_counts["pairs"] if rep_counts else 0, contacts)
    sub_gbs = value * sb["bytes_per_vertex_substep"] / 1e9

    # plain path on the same workload
    plain = None
    if not a.no_plain and garment:
        f0 = a.warmup + a.steps + 1
        solver.SkinObstacles([T[f0]])
        solver.FramesPlain(1)
        solver.sync()
        t1 = time.perf_counter()
        for f in range(a.plain_frames):
            solver.SkinObstacles([T[f0 + 1 + f]])
            solver.FramesPlain(1)
        solver.sync()
        pms = (time.perf_counter() - t1) * 1e3
        plain = n_verts * params.num_substeps * a.plain_frames / (pms * 1e-3)
        print(f"[bench] plain {pms:.1f} ms / {a.plain_frames} frame(s)", file=sys.stderr,
              flush=True)
    elif not a.no_plain:
        solver.FramesPlain(1)
        solver.sync()
        pms = solver.time_frames(False, a.plain_frames)
        plain = n_verts * params.num_substeps * a.plain_frames / (pms * 1e-3)
        print(f"[bench] plain {pms:.1f} ms / {a.plain_frames} frame(s)", file=sys.stderr,
              flush=True)
    rollout = None
    if a.workload == "batched" and a.rollout_frames > 0 and not a.no_graph:
        plain_at = set() if a.no_plain else {int(f) for f in a.plain_at.split(",") if f.strip()}
        print(f"[bench] rollout of {a.rollout_frames} frames (plain at {sorted(plain_at)})",
              file=sys.stderr, flush=True)
        rollout = rollout_leg(solver, insts, specs, a.rollout_frames, plain_at)
        print(f"[bench] rollout: batched {rollout['batched_value']:.4e} vs/s, "
              f"ratio {rollout['batched_over_plain']}", file=sys.stderr, flush=True)
    solver.close()

    cpu, cpu_par, parity, parity_n = None, None, None, 0
    if not a.no_cpu:
        tw_setup = None
        if bound:
            def tw_setup(k, twin):
                twin.setup_bindings(gs[k][2], gs[k][3], gs[k][4], params.bvh_tolerance)
        elif quality:
            from oracle import oracle as orc

            def tw_setup(k, twin):
                # the solver colors faces like every constraint set: seed + instance
                col, ncol = orc.color(specs[k].tris, len(specs[k].pos), a.seed + k,
                                      params.coloring)
                twin.setup_fem(col, ncol, specs[k].pos,
                               (params.fem_xx_stiffness, params.fem_yy_stiffness,
                                params.fem_xy_stiffness, params.fem_xy_poisson_ratio,
                                params.fem_yx_poisson_ratio), params.fem_bending_stiffness)
        cpu, cpu_par, parity, parity_n = cpu_baseline_and_parity(
            specs, params, a.seed, contacts, a.cpu_budget, before, tw_setup,
            parallel=not a.no_cpu_parallel)
        print(f"[bench] cpu {cpu['value']:.3e} vs/s (1 core)"
              + (f", {cpu_par['value']:.3e} vs/s ({cpu_par['cores']} cores)" if cpu_par else "")
              + f", parity {parity:.3e} over {parity_n} instances", file=sys.stderr, flush=True)

    adj = None
    if not a.no_adjoint:
        print("[bench] differentiable rollout", file=sys.stderr, flush=True)
        adj = adjoint_leg()
        print(f"[bench] adjoint substep ratio {adj['substep_ratio']:.2f}, rollout ratio "
              f"{adj['rollout_ratio']:.2f}", file=sys.stderr, flush=True)

    out = {
        "metric": "vertex-substeps/s batched vs plain on 1 GPU at fixed err; adjoint:forward ratio",
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wname, "instances": len(specs), "vertices": n_verts,
                   "substeps": params.num_substeps, "iterations": params.num_iterations,
                   "bend_model": a.bend_model, "obstacle_contact": contacts,
                   "self_contact": self_contact,
                   "particle_collision": swift, "bindings": bound,
                   "constraints": "FEM strain + isometric bending" if quality else "XPBD",
                   "skinning": "GPU LBS every frame" if garment else None,
                   "step": "one frame of all instances",
                   "graph": not a.no_graph},
        "plain_value": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "parity_rel_err": parity,
        "parity_instances": parity_n,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": ("profiles/traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                        "of the same launches (own runs; same algorithmic bytes)")
                     if traffic else None,
                     "kernel": kernel, "kernel_ms": kern_ms, "bytes_per_launch": bytes_kern,
                     "repulsion_counts": rep_counts, "particle": particle,
                     "binding": binding, "fem": fem},
        "roofline_substep": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                             "achieved": sub_gbs, "frac": sub_gbs / HBM_PEAK_GBS,
                             "ceiling_vertex_substeps_per_s":
                                 HBM_PEAK_GBS * 1e9 / sb["bytes_per_vertex_substep"],
                             **sb},
        "rollout": rollout,
        "tuning": tuning,
        "record": record,
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_par,
        "adjoint_forward_ratio": adj["substep_ratio"] if adj else None,
        "adjoint": adj,
        "device": info,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
