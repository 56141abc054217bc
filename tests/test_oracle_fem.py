"""The oracle's quality-mode FEM constraints (oracle/fem_oracle.c), pinned without the code
under test: strain energy zero at rest and under rigid motions, its constraint gradient
equal to central differences of the energy (reciprocal orthotropic parameters, where the
energy 0.5 e:Ce has Ce as its stress), the closed-form energy of uniaxial stretch; the
isometric bending weights' cotangent identity sum_i K_i x_i = 0 on unfolded planar hinges,
the energy 0.5 coeff |sum_i K_i x_i|^2 and its gradient, and an XPBD step that lowers it."""
import numpy as np
import pytest

from oracle import fem

T = np.array([0, 1, 2], np.int32)
B = np.array([0, 1, 2, 3], np.int32)


def _fd(fn, x, h=1e-6):
    g = np.zeros_like(x)
    for v in range(x.shape[0]):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, c] += h
            xm[v, c] -= h
            g[v, c] = (fn(xp) - fn(xm)) / (2 * h)
    return g


def _rigid(rng):
    R, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    return R, rng.standard_normal(3)


@pytest.mark.parametrize("E", [[1.0, 1.0, 0.7, 0.3, 0.3], [2.0, 1.0, 0.5, 0.4, 0.2]])
def test_strain_rest_rigid_and_gradient(E):
    rng = np.random.default_rng(0)
    E = np.asarray(E)
    for _ in range(5):
        x = rng.standard_normal((3, 3))
        area, inv = fem.strain_init(x, T[None])
        assert abs(area[0] - 0.5 * np.linalg.norm(np.cross(x[1] - x[0], x[2] - x[0]))) < 1e-12
        assert abs(fem.strain_energy(x, T, area[0], inv[0], E)[0]) < 1e-25
        R, t = _rigid(rng)
        assert abs(fem.strain_energy(x @ R.T + t, T, area[0], inv[0], E)[0]) < 1e-24
        xd = x + 0.1 * rng.standard_normal((3, 3))
        e, g = fem.strain_energy(xd, T, area[0], inv[0], E)
        fd = _fd(lambda y: fem.strain_energy(y, T, area[0], inv[0], E)[0], xd)
        np.testing.assert_allclose(g, fd, atol=1e-7 * np.abs(fd).max())
        np.testing.assert_allclose(g.sum(0), 0.0, atol=1e-12)  # translation invariance


def test_strain_uniaxial_closed_form():
    # unit right triangle in the xy plane stretched by s along the rest frame's first axis
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    area, inv = fem.strain_init(x, T[None])
    Ex, Ey, G = 3.0, 2.0, 0.5
    E = np.array([Ex, Ey, G, 0.0, 0.0])
    s = 1.2
    for axis in (0, 1):
        y = x.copy()
        y[:, axis] *= s
        e, _ = fem.strain_energy(y, T, area[0], inv[0], E)
        eps = 0.5 * (s * s - 1.0)
        # one of the two material axes carries the strain, no Poisson coupling
        assert min(abs(e - 0.5 * 0.5 * Ex * eps ** 2), abs(e - 0.5 * 0.5 * Ey * eps ** 2)) < 1e-12


def _hinge(rng):
    a, b = rng.standard_normal(2), rng.standard_normal(2)
    d = b - a
    nrm = np.array([-d[1], d[0]])
    c2 = a + rng.uniform(0.1, 0.9) * d + rng.uniform(0.2, 1.0) * nrm
    c3 = a + rng.uniform(0.1, 0.9) * d - rng.uniform(0.2, 1.0) * nrm
    P = np.zeros((4, 3))
    P[:, :2] = [a, b, c2, c3]
    R, t = _rigid(rng)
    return P @ R.T + t


def test_isometric_weights_energy_gradient_and_step():
    rng = np.random.default_rng(1)
    for _ in range(10):
        P = _hinge(rng)
        K, c = fem.iso_init(P, B[None])
        assert np.abs(K[0] @ P).max() < 1e-12      # flat hinge: no bending
        assert abs(K[0].sum()) < 1e-12              # translation invariance
        Q = P.copy()
        Q[2] += 0.2 * rng.standard_normal(3)        # fold the hinge
        e, g = fem.iso_energy(Q, B, K[0], c[0])
        assert abs(e - 0.5 * c[0] * np.sum((K[0] @ Q) ** 2)) < 1e-10 * max(e, 1.0)
        fd = _fd(lambda y: fem.iso_energy(y, B, K[0], c[0])[0], Q)
        np.testing.assert_allclose(g, fd, atol=1e-6 * np.abs(fd).max())
        # one XPBD step at very large stiffness lowers the energy
        x = np.ascontiguousarray(Q.reshape(-1))
        lam = np.zeros(1)
        fem.iso_solve(x.reshape(-1, 3), np.ones(4), B[None], K, c, lam, 0, 1, 0, 1e-3, 1e12)
        assert fem.iso_energy(x.reshape(-1, 3), B, K[0], c[0])[0] < e


def _flat_sheet(n=10, side=0.3, seed=2):
    from xrtailor_amd import scenes
    pos, tris = scenes.grid(n, side, 0.0)
    rng = np.random.default_rng(seed)
    h = side / (n - 1)
    pos[:, [0, 2]] += 0.25 * h * rng.uniform(-1.0, 1.0, (len(pos), 2))  # in plane only
    return pos, tris


def _fem_twin(pos, tris, x0, E, **kw):
    from contact_helpers import oracle_scene
    from oracle import oracle as orc
    from xrtailor_amd import scenes
    from xrtailor_amd.solver import make_params
    xp = make_params(constraint_model=1, gravity=(0.0, 0.0, 0.0), enable_ground=0, **kw)
    spec = scenes.ClothSpec(pos, tris, np.zeros(0, np.int32), 1e-7, 1e-3, x0.astype(np.float32))
    twin = oracle_scene(spec, xp, seed=0)
    col, ncol = orc.color(tris, len(pos), 0, 0)
    twin.setup_fem(col, ncol, pos, E, 1.0)
    return twin


def test_scene_fem_rest_under_rigid_motion_is_a_fixed_point():
    """The twin's FEM iteration (face colors, strain and isometric bending wiring) leaves a
    flat sheet at rest, moved rigidly, exactly in place without gravity."""
    pos, tris = _flat_sheet()
    R, t = _rigid(np.random.default_rng(3))
    x0 = (pos @ R.T + t).astype(np.float32)
    twin = _fem_twin(pos, tris, x0, (1.0, 1.0, 1.0, 0.3, 0.3))
    start = twin.c.state()[0].copy()
    for _ in range(3):
        twin.substep()
    np.testing.assert_allclose(twin.c.state()[0], start, rtol=0, atol=1e-9)


def test_scene_fem_strain_relaxes_a_stretched_sheet():
    """A sheet stretched by 10 % along x (moduli large enough for the projection to act)
    contracts toward its rest lengths under the twin's FEM iterations (mean edge strain;
    single edges overshoot while the velocities the projections create carry over)."""
    pos, tris = _flat_sheet()
    x0 = pos * np.array([1.1, 1.0, 1.0])
    twin = _fem_twin(pos, tris, x0, (1e3, 1e3, 4e2, 0.3, 0.3), solve_bending=0)

    def strain(x):
        e = np.unique(np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), axis=1), axis=0)
        L0 = np.linalg.norm(pos[e[:, 0]] - pos[e[:, 1]], axis=1)
        return np.abs(np.linalg.norm(x[e[:, 0]] - x[e[:, 1]], axis=1) / L0 - 1.0).mean()

    s0 = strain(twin.c.state()[0])
    for _ in range(2):
        twin.substep()
    s1 = strain(twin.c.state()[0])
    assert s0 > 0.045 and s1 < 0.5 * s0
