"""Procedural scenes of the benchmark / test workloads and their seeded RNG.

RNG: SplitMix64 streams with explicit Box-Muller normals, so that every workload is
reproducible from its seed on any host (numpy's generators are not used).

Workloads (sizes follow the solver's benchmark configurations):
  hanging_cloth   64x64 grid, 1 m, two top corners pinned, stretch 1e-8 / bend 1e-4
  drop_cloth      256x256 grid, 2 m, 0.5 m above an icosphere (s5, r 0.5)
  batched         256 instances, side round(32*6^u), log-uniform compliances
  garment         tube garment (top ring = LRA slots) around a skinned capsule-chain body
  diff_rollout    128x128 cloth over a sphere
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """SplitMix64 stream: value k is mix(seed + (k+1) * golden)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def uint64(self, n: int) -> np.ndarray:
        k = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            return _mix(self.state + k * _GOLDEN)

    def uniform(self, n: int) -> np.ndarray:
        """[0, 1) doubles from the top 53 bits."""
        return (self.uint64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller: u1 in (0,1], u2 in [0,1)."""
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


# --------------------------------------------------------------------- meshes
def grid(n: int, side: float, height: float = 0.0):
    """n x n vertex grid in the y=height plane, x,z in [-side/2, side/2].
    Vertex (i, j) has index j*n + i (i along x, j along z).  Quads are split on the
    (i,j)-(i+1,j+1) diagonal, wound counter-clockwise seen from +y."""
    t = np.linspace(-0.5 * side, 0.5 * side, n)
    xs, zs = np.meshgrid(t, t, indexing="xy")
    pos = np.stack([xs.ravel(), np.full(n * n, height), zs.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (j * n + i).ravel()
    b = a + 1
    c = a + n + 1
    d = a + n
    tris = np.empty((2 * a.size, 3), np.int32)
    tris[0::2] = np.stack([a, c, b], 1)
    tris[1::2] = np.stack([a, d, c], 1)
    return pos, tris


def icosphere(subdiv: int, radius: float, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron (10*4^s + 2 vertices), outward CCW winding."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache: dict = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    pos = np.array(verts) * radius + np.asarray(center, float)
    return pos, np.array(faces, np.int32)


def capsule(radius: float, half_length: float, subdiv: int = 3, center=(0.0, 0.0, 0.0)):
    """Capsule along x: an icosphere whose +x / -x halves are moved apart."""
    pos, tris = icosphere(subdiv, radius)
    pos[:, 0] += np.sign(pos[:, 0]) * half_length * (np.abs(pos[:, 0]) > 1e-9)
    return pos + np.asarray(center, float), tris


def tube(ring: int, rows: int, radius: float, height: float, y0: float = 0.0):
    """Open cylinder around the y axis: `rows` rings of `ring` vertices, ring j at height
    y0 + height * j / (rows - 1); vertex (i, j) has index j*ring + i.  Quads are split on
    the (i,j)-(i+1,j+1) diagonal, wound counter-clockwise seen from outside."""
    th = 2.0 * np.pi * np.arange(ring) / ring
    ys = y0 + height * np.arange(rows) / (rows - 1)
    pos = np.empty((rows * ring, 3))
    for j in range(rows):
        pos[j * ring:(j + 1) * ring] = np.stack(
            [radius * np.cos(th), np.full(ring, ys[j]), radius * np.sin(th)], 1)
    tris = []
    for j in range(rows - 1):
        for i in range(ring):
            a, b = j * ring + i, j * ring + (i + 1) % ring
            c, d = (j + 1) * ring + (i + 1) % ring, (j + 1) * ring + i
            tris += [(a, c, b), (a, d, c)]
    return pos, np.array(tris, np.int32)


def capsule_y(radius: float, half_length: float, subdiv: int = 3, center=(0.0, 0.0, 0.0)):
    """Capsule along y: an icosphere whose +y / -y halves are moved apart."""
    pos, tris = icosphere(subdiv, radius)
    pos[:, 1] += np.sign(pos[:, 1]) * half_length * (np.abs(pos[:, 1]) > 1e-9)
    return pos + np.asarray(center, float), tris


def capsule_latlong_y(radius: float, half_length: float, ring: int = 56, rows: int = 123,
                      center=(0.0, 0.0, 0.0)):
    """Capsule along y as a latitude-longitude mesh: a pole vertex at each end and `rows`
    rings of `ring` vertices spaced evenly in arc length along the profile (quarter circle,
    straight side, quarter circle).  ring * rows + 2 vertices, 2 ring * rows triangles
    (defaults: 6,890 vertices and 13,776 triangles, the vertex and face counts of an SMPL
    body mesh), wound counter-clockwise seen from outside.  Vertex 0 is the bottom pole,
    ring j (from the bottom) holds 1 + j ring + i, the last vertex is the top pole."""
    arc = 0.5 * np.pi * radius
    total = 2.0 * arc + 2.0 * half_length
    s = total * np.arange(1, rows + 1) / (rows + 1)
    rho = np.empty(rows)
    y = np.empty(rows)
    lo, hi = s < arc, s > arc + 2.0 * half_length
    mid = ~(lo | hi)
    rho[lo] = radius * np.sin(s[lo] / radius)
    y[lo] = -half_length - radius * np.cos(s[lo] / radius)
    rho[mid] = radius
    y[mid] = -half_length + (s[mid] - arc)
    rho[hi] = radius * np.sin((total - s[hi]) / radius)
    y[hi] = half_length + radius * np.cos((total - s[hi]) / radius)
    th = 2.0 * np.pi * np.arange(ring) / ring
    pos = np.empty((ring * rows + 2, 3))
    pos[0] = (0.0, -half_length - radius, 0.0)
    pos[-1] = (0.0, half_length + radius, 0.0)
    for j in range(rows):
        pos[1 + j * ring:1 + (j + 1) * ring] = np.stack(
            [rho[j] * np.cos(th), np.full(ring, y[j]), rho[j] * np.sin(th)], 1)
    top = ring * rows + 1
    tris = []
    for i in range(ring):
        i1 = (i + 1) % ring
        tris.append((0, 1 + i, 1 + i1))
    for j in range(rows - 1):
        for i in range(ring):
            i1 = (i + 1) % ring
            a, b = 1 + j * ring + i, 1 + j * ring + i1
            c, d = 1 + (j + 1) * ring + i1, 1 + (j + 1) * ring + i
            tris += [(a, c, b), (a, d, c)]
    for i in range(ring):
        i1 = (i + 1) % ring
        base = 1 + (rows - 1) * ring
        tris.append((top, base + i1, base + i))
    return pos + np.asarray(center, float), np.array(tris, np.int32)


@dataclass
class ChainBody:
    """A capsule-chain body: one capsule along y over `joints` segments of length `seg`
    starting at y0, skinned to the chain by linear blend skinning.  Joint j bends the
    chain above it about the x and z axes by theta_j(t) = sum_k a_jk (sin(2 pi f_jk t +
    phi_jk) - sin(phi_jk)) (three terms per axis, f in [0.2, 2] Hz, sum |a| <= 0.6 rad per
    joint; the offset starts the motion from the rest pose)."""
    pos: np.ndarray    # rest vertices (nb, 3)
    tris: np.ndarray
    y0: float
    seg: float
    amp: np.ndarray    # (joints, 2, 3) amplitudes, frequencies, phases per axis and term
    freq: np.ndarray
    phase: np.ndarray

    def angles(self, t: float) -> np.ndarray:
        """(joints, 2) bend angles about x and z, zero at t = 0 (the rest pose)."""
        a = self.amp * (np.sin(2.0 * np.pi * self.freq * t + self.phase) - np.sin(self.phase))
        return a.sum(-1)

    def _kinematics(self, t: float):
        """Per joint: rotation and (rest origin, current origin) at time t."""
        nj = self.amp.shape[0]
        ang = self.angles(t)
        rot = np.eye(3)
        rots, origins = [], []
        p_rest = np.array([0.0, self.y0, 0.0])
        p_now = p_rest.copy()
        for j in range(nj):
            ax, az = ang[j]
            cx, sx, cz, sz = np.cos(ax), np.sin(ax), np.cos(az), np.sin(az)
            rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
            rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
            rot = rot @ rx @ rz
            rots.append(rot.copy())
            origins.append((p_rest.copy(), p_now.copy()))
            step = np.array([0.0, self.seg, 0.0])
            p_rest = p_rest + step
            p_now = p_now + rot @ step
        return rots, origins

    def skinning(self):
        """Sparse LBS weights of the rest vertices as CSR (ptr, joint, weight): the blend
        of at()."""
        nj = self.amp.shape[0]
        u = np.clip((self.pos[:, 1] - self.y0) / self.seg, 0.0, nj - 1e-9)
        s = np.floor(u).astype(int)
        f = u - s
        w_lo = np.where((f < 0.5) & (s > 0), 0.5 - f, 0.0)
        ptr, joint, weight = [0], [], []
        for v in range(len(self.pos)):
            if w_lo[v] > 0.0:
                joint += [s[v] - 1, s[v]]
                weight += [w_lo[v], 1.0 - w_lo[v]]
            else:
                joint.append(s[v])
                weight.append(1.0)
            ptr.append(len(joint))
        return (np.asarray(ptr, np.int32), np.asarray(joint, np.int32),
                np.asarray(weight, np.float32))

    def transforms(self, t: float) -> np.ndarray:
        """Joint transforms at time t, (joints, 12) float32 row-major 3x4 [R | r1 - R r0]:
        the skinned position of a vertex is sum_k w_k (R_k p + t_k)."""
        rots, origins = self._kinematics(t)
        out = np.zeros((len(rots), 12))
        for k, (R, (r0, r1)) in enumerate(zip(rots, origins)):
            out[k] = np.concatenate([R, (r1 - R @ r0)[:, None]], axis=1).reshape(-1)
        return out.astype(np.float32)

    def at(self, t: float) -> np.ndarray:
        """Skinned vertex positions at time t (float32)."""
        nj = self.amp.shape[0]
        ang = self.angles(t)
        rot = np.eye(3)
        rots, origins = [], []
        p_rest = np.array([0.0, self.y0, 0.0])
        p_now = p_rest.copy()
        for j in range(nj):  # forward kinematics: joint j at y0 + j seg
            ax, az = ang[j]
            cx, sx, cz, sz = np.cos(ax), np.sin(ax), np.cos(az), np.sin(az)
            rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
            rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
            rot = rot @ rx @ rz
            rots.append(rot.copy())
            origins.append((p_rest.copy(), p_now.copy()))
            step = np.array([0.0, self.seg, 0.0])
            p_rest = p_rest + step
            p_now = p_now + rot @ step
        # segment s covers [y0 + s seg, y0 + (s+1) seg); blend linearly with the segment
        # below over the lower half of each segment
        u = np.clip((self.pos[:, 1] - self.y0) / self.seg, 0.0, nj - 1e-9)
        s = np.floor(u).astype(int)
        f = u - s
        w_lo = np.where((f < 0.5) & (s > 0), 0.5 - f, 0.0)
        out = np.zeros_like(self.pos)
        for k in range(nj):
            wk = np.where(s == k, 1.0 - w_lo, 0.0) + np.where(s == k + 1, w_lo, 0.0)
            if not wk.any():
                continue
            r0, r1 = origins[k]
            moved = (self.pos - r0) @ rots[k].T + r1
            out += wk[:, None] * moved
        return out.astype(np.float32)


def chain_body(seed: int, joints: int, radius: float, seg: float, y0: float,
               subdiv: int = 3, mesh: str = "ico", amp_total: float | None = None) -> ChainBody:
    """mesh "ico": icosphere-based capsule of subdivision `subdiv`; "latlong": the 6,890-
    vertex latitude-longitude capsule (capsule_latlong_y).  Each joint's amplitudes sum to
    0.6 rad, or, with amp_total, to amp_total / joints (the chain's compound bend stays
    within amp_total: 24 joints of 0.6 rad each would swing the top of a 0.8 m chain by
    metres)."""
    rng = Rng(seed)
    half = 0.5 * joints * seg - radius
    ctr = (0.0, y0 + 0.5 * joints * seg, 0.0)
    if mesh == "latlong":
        pos, tris = capsule_latlong_y(radius, max(half, 0.0), center=ctr)
    else:
        pos, tris = capsule_y(radius, max(half, 0.0), subdiv, ctr)
    amp = rng.uniform(joints * 2 * 3).reshape(joints, 2, 3)
    per_joint = 0.6 if amp_total is None else min(0.6, amp_total / joints)
    amp = per_joint * amp / amp.sum(-1, keepdims=True).sum(-2, keepdims=True)  # sum |a| per joint
    freq = 0.2 + 1.8 * rng.uniform(joints * 2 * 3).reshape(joints, 2, 3)
    phase = 2.0 * np.pi * rng.uniform(joints * 2 * 3).reshape(joints, 2, 3)
    return ChainBody(pos, tris, y0, seg, amp, freq, phase)


# ------------------------------------------------------------------ workloads
@dataclass
class ClothSpec:
    pos: np.ndarray            # rest positions (nv, 3) float64
    tris: np.ndarray           # (nt, 3) int32
    pinned: np.ndarray         # pinned vertex ids
    stretch_compliance: float
    bend_compliance: float
    x0: np.ndarray             # initial (jittered) positions (nv, 3) float32
    obstacle: tuple | None = None   # (pos, tris) or None
    lra_slots: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))


def hanging_cloth(seed: int = 1, n: int = 64, side: float = 1.0, jitter: float = 1e-4,
                  stretch: float = 1e-8, bend: float = 1e-4) -> ClothSpec:
    pos, tris = grid(n, side, 0.0)
    pinned = np.array([0, n - 1], np.int32)  # the two corners of row z = -side/2
    x0 = (pos + jitter * Rng(seed).normal(pos.size).reshape(-1, 3)).astype(np.float32)
    return ClothSpec(pos, tris, pinned, stretch, bend, x0)


def drop_cloth(seed: int = 2, n: int = 256, side: float = 2.0, jitter: float = 1e-3,
               height: float = 1.0) -> ClothSpec:
    pos, tris = grid(n, side, height)
    x0 = (pos + jitter * Rng(seed).normal(pos.size).reshape(-1, 3)).astype(np.float32)
    obstacle = icosphere(5, 0.5)
    return ClothSpec(pos, tris, np.zeros(0, np.int32), 1e-8, 1e-4, x0, obstacle)


def differentiable_drop(seed: int = 5, n: int = 128, side: float = 1.0, height: float = 0.35,
                        radius: float = 0.3, jitter: float = 1e-3):
    """Differentiable-rollout workload: an n x n cloth released above a sphere (icosphere
    s4, radius 0.3 m, at the origin), stretch compliance 1e-8 per edge, bend 1e-4."""
    pos, tris = grid(n, side, height)
    rng = Rng(seed)
    x0 = (pos + jitter * rng.normal(pos.size).reshape(-1, 3)).astype(np.float32)
    spec = ClothSpec(pos, tris, np.zeros(0, np.int32), 1e-8, 1e-4, x0, icosphere(4, radius))
    return spec


def garment(seed: int = 6, ring: int = 160, rows: int = 125, joints: int = 24,
            radius: float = 0.2, height: float = 0.6, gap: float = 0.01,
            jitter: float = 1e-3, body_subdiv: int = 5, body_mesh: str = "ico",
            amp_total: float | None = None) -> tuple[ClothSpec, ChainBody]:
    """Tube garment (ring x rows vertices, radius `radius`, top ring at y = height as LRA
    slots) around a capsule-chain body whose surface lies `gap` inside it; stretch 1e-8,
    bend 1e-4.  The body moves by ChainBody.at(t)."""
    pos, tris = tube(ring, rows, radius, height, 0.0)
    rng = Rng(seed)
    x0 = (pos + jitter * rng.normal(pos.size).reshape(-1, 3)).astype(np.float32)
    seg = (height + 0.2) / joints
    body = chain_body(seed + 1, joints, radius - gap, seg, -0.1, body_subdiv, body_mesh, amp_total)
    slots = np.arange((rows - 1) * ring, rows * ring, dtype=np.int32)
    spec = ClothSpec(pos, tris, np.zeros(0, np.int32), 1e-8, 1e-4, x0, (body.pos, body.tris),
                     slots)
    return spec, body


def garment_on_body(seed: int = 6, gap: float = 0.01):
    """The garment-on-animated-body workload at its stated shape: a 160 x 125 tube garment
    (20,000 vertices, radius 0.2 m, 0.6 m tall) whose top (shoulder) ring holds the 160
    long-range attachment slots (every slot to every vertex) and is bound to the body
    (stiffness 1, at its start distance), around a capsule-chain body of 6,890 vertices
    (latitude-longitude capsule) skinned by 24 joints, `gap` inside the tube; joint angles
    are seeded sums of three sinusoids (0.2-2 Hz) whose amplitudes sum to 0.6 rad over the
    chain.  Returns (spec, body, bound vertices, stiffness, distance)."""
    ring, rows = 160, 125
    spec, body = garment(seed=seed, ring=ring, rows=rows, joints=24, gap=gap, body_mesh="latlong",
                         amp_total=0.6)
    top = np.arange((rows - 1) * ring, rows * ring, dtype=np.int32)
    return spec, body, top, np.ones(ring, np.float32), np.zeros(ring, np.float32)


def garments(seed: int = 4, n_instances: int = 64, ring_lo: int = 64, ring_hi: int = 128,
             rows_lo: int = 40, rows_hi: int = 80, joints: int = 12, body_subdiv: int = 5):
    """Garment instances for the pipeline around the substep: tube garments of uneven size
    (ring and rows uniform in the given ranges) around capsule-chain bodies of 10,242
    vertices (subdivision 5) skinned by `joints` joints.  Every 8th vertex of the top ring
    is an LRA slot; the top ring is bound to the body with stiffness 1 and the middle row
    with stiffness 0.5, both at their distance at the start.  Returns (spec, body, bound
    vertices, stiffness, distance) per instance."""
    rng = Rng(seed)
    rings = np.rint(ring_lo + (ring_hi - ring_lo) * rng.uniform(n_instances)).astype(int)
    rows = np.rint(rows_lo + (rows_hi - rows_lo) * rng.uniform(n_instances)).astype(int)
    out = []
    for i in range(n_instances):
        r, h = int(rings[i]), int(rows[i])
        spec, body = garment(seed=seed * 7919 + i, ring=r, rows=h, joints=joints,
                             body_subdiv=body_subdiv)
        spec.lra_slots = spec.lra_slots[::8].copy()
        top = np.arange((h - 1) * r, h * r, dtype=np.int32)
        mid = np.arange((h // 2) * r, (h // 2 + 1) * r, dtype=np.int32)
        bind = np.concatenate([top, mid])
        stiff = np.concatenate([np.ones(r), np.full(r, 0.5)]).astype(np.float32)
        out.append((spec, body, bind, stiff, np.zeros(2 * r, np.float32)))
    return out


def batched(seed: int = 4, n_instances: int = 256, side_lo: int = 32, side_hi: int = 192,
            cloth_side: float = 1.0, jitter: float = 1e-3) -> list[ClothSpec]:
    """Independent instances: grid side round(lo*(hi/lo)^u) (log-uniform), stretch
    compliance 10^U(-9,-6), bend 10^U(-5,-3), sphere or capsule obstacle r~U(0.2,0.6)
    below the cloth."""
    rng = Rng(seed)
    u = rng.uniform(n_instances)
    sides = np.rint(side_lo * (side_hi / side_lo) ** u).astype(int)
    sc = 10.0 ** (-9.0 + 3.0 * rng.uniform(n_instances))
    bc = 10.0 ** (-5.0 + 2.0 * rng.uniform(n_instances))
    kind = rng.uniform(n_instances) < 0.5
    radius = 0.2 + 0.4 * rng.uniform(n_instances)
    out = []
    for i in range(n_instances):
        pos, tris = grid(int(sides[i]), cloth_side, 0.0)
        x0 = (pos + jitter * Rng(seed * 1000003 + i).normal(pos.size).reshape(-1, 3)).astype(np.float32)
        r = float(radius[i])
        ctr = (0.0, -r - 0.05, 0.0)
        obst = icosphere(3, r, ctr) if kind[i] else capsule(r, 0.5 * r, 3, ctr)
        out.append(ClothSpec(pos, tris, np.zeros(0, np.int32), float(sc[i]), float(bc[i]), x0, obst))
    return out
