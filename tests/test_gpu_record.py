"""GPU per-frame cache (xrt_record_frames): the recorded positions equal the state after
each frame bit for bit, the recorded normals equal the FP32 oracle's vertex normals
(UpdateNormals' formula) of those positions bit for bit, for instances of uneven size."""
import numpy as np
import pytest

from gpu_helpers import build
from oracle import oracle as orc
from xrtailor_amd import scenes
from xrtailor_amd.solver import make_params

pytestmark = pytest.mark.gpu


def _specs():
    return [scenes.hanging_cloth(seed=1, n=24), scenes.hanging_cloth(seed=2, n=40),
            scenes.hanging_cloth(seed=3, n=17)]


@pytest.mark.parametrize("with_normals", [True, False])
def test_record_frames_match_state_and_oracle_normals(with_normals):
    print(f"case record frames, normals {with_normals}", flush=True)
    specs = _specs()
    xp = make_params(bend_model=1, num_substeps=4, num_iterations=4)
    s, insts = build(specs, xp, seed=1)
    pos, nrm = s.RecordFrames(3, normals=with_normals)
    ref, _ = build(specs, xp, seed=1)
    offs = np.cumsum([0] + [len(sp.pos) for sp in specs])
    for f in range(3):
        ref.FramesBatched(1)
        for k, i in enumerate(insts):
            x = ref.GetState(i)[0]
            np.testing.assert_array_equal(pos[f, offs[k]:offs[k + 1]], x)
            if with_normals:
                want = orc.vertex_normals(x, specs[k].tris)
                np.testing.assert_array_equal(nrm[f, offs[k]:offs[k + 1]], want)
    assert nrm is None or np.isfinite(nrm).all()
    # a second call continues from the state (the slots are reused)
    pos2, _ = s.RecordFrames(2, normals=with_normals)
    ref.FramesBatched(1)
    np.testing.assert_array_equal(pos2[0, offs[1]:offs[2]], ref.GetState(insts[1])[0])
    print(f"  {pos.shape[1]} vertices x 3 frames recorded", flush=True)
    s.close()
    ref.close()
