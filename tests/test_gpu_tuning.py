"""The autotuner and the launch shapes it chooses from: every candidate shape gives the
batched substep's results bit for bit, the autotuner leaves the state as it found it, and
a tuned context's frames equal an untuned one's (persistent and wide paths)."""
import numpy as np
import pytest

from xrtailor_amd import XrtError, scenes
from xrtailor_amd.solver import build_solver, make_params

pytestmark = pytest.mark.gpu


def _specs(n=12):
    # sizes from the batched workload's distribution, including instances too large for LDS
    return scenes.batched(seed=4, n_instances=256)[::21][:n]


def _states(s, ids):
    return [s.GetState(i) for i in ids]


def _equal(a, b):
    for (xa, va), (xb, vb) in zip(a, b):
        np.testing.assert_array_equal(xa, xb)
        np.testing.assert_array_equal(va, vb)


def test_persistent_candidates_equal_default_and_autotune_restores_state():
    print("case autotune: persistent substep candidates (block x unroll) and the wide substep",
          flush=True)
    specs = _specs()
    print(f"  instances of {sorted(len(sp.pos) for sp in specs)} vertices", flush=True)
    xp = make_params(bend_model=1, ground_height=-1.0)
    s, ids = build_solver(specs, xp, seed=4)
    start = [(sp.x0, np.zeros_like(sp.x0)) for sp in specs]
    ref = None
    shapes = [(block, unroll, 1, 0) for block in (1024, 768, 512, 256) for unroll in (0, 1, 2)]
    shapes += [(1024, 0, 1, 1), (1024, 0, 2, 1)]  # wide substep, 1 / 2 elements per thread
    for shape in shapes:
        for i, (x0, v0) in zip(ids, start):
            s.SetState(i, x0, v0)
        s.set_tuning(*shape)
        s.FramesBatched(1)
        s.sync()
        got = _states(s, ids)
        if ref is None:
            ref = got
        else:
            _equal(got, ref)
    before = _states(s, ids)
    t = s.autotune(2)
    print(f"  chose {t}", flush=True)
    assert t["path"] in (0, 1) and t["candidates"] == 14 and t["best_ms"] > 0.0
    assert t["best_ms"] <= t["default_ms"]
    _equal(_states(s, ids), before)                 # state restored
    assert s.tuning()["persistent_block"] == t["persistent_block"]
    with pytest.raises(XrtError):
        s.set_tuning(100, 0, 1)
    with pytest.raises(XrtError):
        s.set_tuning(1024, 3, 1)
    with pytest.raises(XrtError):
        s.set_tuning(1024, 0, 1, 2)


def test_tuned_frames_equal_untuned_frames():
    print("case autotune: tuned vs untuned frames, persistent and wide (contacts) paths",
          flush=True)
    specs = _specs(8)
    for contacts in (0, 1):
        xp = make_params(bend_model=1, ground_height=-1.0, enable_obstacle_contact=contacts,
                         enable_self_contact=contacts)
        runs = []
        for mode in ("tuned", "default", "two per thread"):
            s, ids = build_solver(specs, xp, seed=4)
            if mode == "tuned":
                t = s.autotune(2)
                print(f"  contacts={contacts}: {t}", flush=True)
                assert t["candidates"] == (2 if contacts else 14)
                assert t["path"] == 1 if contacts else t["path"] in (0, 1)
            elif mode == "two per thread":
                s.set_tuning(1024, 0, 2)
            s.FramesBatched(2)
            s.sync()
            runs.append(_states(s, ids))
            s.close()
        _equal(runs[0], runs[1])
        _equal(runs[2], runs[1])


def test_autotune_between_substeps_of_a_frame_changes_nothing():
    """The autotuner run between the substeps of a frame (here after one frame and three
    single substeps, with body and self contact: contact records and the frame schedule in
    use) puts back everything later substeps read, so the following frames equal those of a
    context that never ran it."""
    print("case autotune mid-frame with contacts vs no autotune", flush=True)
    specs = _specs(8)
    xp = make_params(bend_model=1, ground_height=-1.0, enable_obstacle_contact=1,
                     enable_self_contact=1, pair_capacity_per_face=128)
    runs = []
    for tune in (True, False):
        s, ids = build_solver(specs, xp, seed=4)
        s.FramesBatched(1)
        for _ in range(3):
            s.SubstepBatched()
        if tune:
            t = s.autotune(3)
            print(f"  chose {t}", flush=True)
            s.set_tuning(1024, 0, 1)     # the defaults, so that only the autotuner differs
        s.FramesBatched(2)
        s.sync()
        runs.append(_states(s, ids))
        s.close()
    _equal(runs[0], runs[1])
