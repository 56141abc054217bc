"""CPU checks of the oracle's contact orchestration: a cloth dropped on a sphere does not
tunnel into it; two sheets closer than the thickness are pushed apart."""
import numpy as np

from contact_helpers import oracle_scene, sphere_drop, two_layers
from xrtailor_amd.solver import make_params


def test_cloth_on_sphere_stays_outside():
    spec, v0 = sphere_drop(n=16)
    xp = make_params(bend_model=1, enable_obstacle_contact=1, enable_self_contact=1,
                     enable_ground=0)
    sc = oracle_scene(spec, xp, v0)
    for _ in range(20):
        sc.substep()
    x0, v = sc.c.state()
    assert np.all(np.isfinite(x0))
    r = np.linalg.norm(x0, axis=1)
    assert r.min() > 0.3 - 2e-3          # nothing deep inside the sphere (radius 0.3)
    near = r < 0.32
    assert near.sum() > 10               # the cloth rests on it


def test_two_layers_separate():
    spec = two_layers()
    xp = make_params(bend_model=1, enable_self_contact=1, enable_ground=0, gravity=(0, 0, 0),
                     pbd_thickness=0.002)
    sc = oracle_scene(spec, xp)
    nv = len(spec.pos) // 2
    gap0 = (spec.x0[nv:, 1] - spec.x0[:nv, 1]).mean()
    for _ in range(5):
        sc.substep()
    assert len(sc.pairs) > 0
    x0, _ = sc.c.state()
    gap = (x0[nv:, 1] - x0[:nv, 1]).mean()
    assert gap > gap0 + 1e-4
