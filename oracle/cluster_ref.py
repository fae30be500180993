"""FP64 numpy restatement of the foam cluster dynamics.

TEST INFRASTRUCTURE ONLY.  Restates ``FoamSimulator`` of
``src/physics/foam_dynamics.rs`` and the parts of ``BubbleCluster`` /
``BubbleConnection`` of ``src/physics/foam.rs`` it drives, in FP64 (the
reference computes in f32 with glam ``Vec3``).  It is the checker of the
HIP cluster kernels (``sf_cluster_*``) and the CPU baseline of the cluster
benchmark; the shipped package never imports it.  Each function names the
reference lines it follows.  Its pins (known answers of the reference's own
tests and closed forms) are listed in ``tests/test_cluster.py``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

AIR_DENSITY = 1.2            # foam_dynamics.rs:106
MAX_SPEED = 0.1              # foam_dynamics.rs:182
ADHESION_STRENGTH = 2.0      # foam_dynamics.rs:147
CONTACT_GAP = 0.3            # foam.rs:342 (30 % of the first bubble's radius)


@dataclass
class Params:
    """FoamSimulator fields and their defaults (foam_dynamics.rs:13-53)."""
    surface_tension: float = 0.025
    viscosity: float = 0.001
    density: float = 1000.0
    gravity: tuple = (0.0, -9.81, 0.0)
    van_der_waals_strength: float = 0.001
    repulsion_stiffness: float = 5.0
    damping: float = 0.8
    coalescence_enabled: bool = False
    coalescence_threshold: float = 0.5


@dataclass
class Cluster:
    """Bubble arrays in vector order (foam.rs:31-49, 214-225)."""
    ids: np.ndarray
    pos: np.ndarray
    radius: np.ndarray
    vel: np.ndarray
    thickness_nm: np.ndarray
    aspect: np.ndarray
    refr: np.ndarray
    next_id: int
    connections: List[tuple] = field(default_factory=list)
    neighbors: List[list] = field(default_factory=list)

    @classmethod
    def from_arrays(cls, pos, radius, vel=None, thickness_nm=None, ids=None):
        n = len(radius)
        ids = np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids, np.int64)
        return cls(ids=ids.copy(), pos=np.array(pos, np.float64).reshape(n, 3),
                   radius=np.array(radius, np.float64),
                   vel=np.zeros((n, 3)) if vel is None else np.array(vel, np.float64).reshape(n, 3),
                   thickness_nm=np.full(n, 500.0) if thickness_nm is None else np.array(thickness_nm, np.float64),
                   aspect=np.ones(n), refr=np.full(n, 1.33),
                   next_id=int(ids.max()) + 1 if n else 0)

    def copy(self):
        return Cluster(self.ids.copy(), self.pos.copy(), self.radius.copy(), self.vel.copy(),
                       self.thickness_nm.copy(), self.aspect.copy(), self.refr.copy(), self.next_id,
                       list(self.connections), [list(x) for x in self.neighbors])

    def __len__(self):
        return len(self.radius)


def default_cluster() -> Cluster:
    """BubbleCluster::create_default_cluster (foam.rs:402-416)."""
    pos = [(0.0, 0.0, 0.0), (0.042, 0.0, 0.0), (-0.038, 0.015, 0.0), (0.0, 0.040, 0.0), (0.018, -0.038, 0.015)]
    c = Cluster.from_arrays(pos, [0.025, 0.02, 0.022, 0.018, 0.02])
    update_connections(c, Params().surface_tension)
    return c


def volume(r):
    """Bubble::volume (foam.rs:66-68)."""
    return (4.0 / 3.0) * math.pi * np.asarray(r) ** 3


def film_mass(c: Cluster, p: Params):
    """Shell volume times density (foam_dynamics.rs:99-102, 171-174)."""
    return 4.0 * math.pi * c.radius ** 2 * (c.thickness_nm * 1e-9) * p.density


def buoyancy(c: Cluster, p: Params) -> np.ndarray:
    """compute_buoyancy (foam_dynamics.rs:96-111): along the gravity
    direction, magnitude max(film weight - displaced air weight, 0)."""
    g = np.asarray(p.gravity, np.float64)
    gl = float(np.linalg.norm(g))
    gdir = g / gl if gl > 0.0 else np.zeros(3)
    net = film_mass(c, p) * gl - volume(c.radius) * AIR_DENSITY * gl
    return gdir[None, :] * np.maximum(net, 0.0)[:, None]


def pair_forces(c: Cluster, p: Params) -> np.ndarray:
    """F[i, j] = force on bubble i from bubble j, compute_interaction_force
    (foam_dynamics.rs:116-154) with a = i, b = j; antisymmetric."""
    delta = c.pos[None, :, :] - c.pos[:, None, :]
    dist = np.sqrt(np.sum(delta * delta, axis=2))
    ok = dist >= 1e-6
    safe = np.where(ok, dist, 1.0)
    direction = delta / safe[:, :, None]
    s = c.radius[:, None] + c.radius[None, :]
    mag = np.zeros_like(dist)
    vdw = (dist < 3.0 * s) & (dist > s)
    mag += np.where(vdw, p.van_der_waals_strength * (c.radius[:, None] * c.radius[None, :]) / safe ** 2, 0.0)
    overlap = s - dist
    mag -= np.where(overlap > 0.0, p.repulsion_stiffness * np.maximum(overlap, 0.0) ** 1.5, 0.0)
    thr = 1.2 * s
    adh = (dist < thr) & (dist > 0.95 * s)
    mag += np.where(adh, ADHESION_STRENGTH * (thr - dist), 0.0)
    mag = np.where(ok, mag, 0.0)
    return direction * mag[:, :, None]


def drag(c: Cluster, p: Params) -> np.ndarray:
    """Stokes drag -6 pi eta R v (foam_dynamics.rs:157-161)."""
    return -c.vel * (6.0 * math.pi * p.viscosity * c.radius)[:, None]


def compute_forces(c: Cluster, p: Params) -> np.ndarray:
    """compute_forces (foam_dynamics.rs:61-92)."""
    return buoyancy(c, p) + pair_forces(c, p).sum(axis=1) + drag(c, p)


def force_scale(c: Cluster, p: Params) -> np.ndarray:
    """Sum of the magnitudes of every term of each bubble's force: the scale
    against which a reordered FP32 sum is compared."""
    return (np.linalg.norm(buoyancy(c, p), axis=1) + np.linalg.norm(pair_forces(c, p), axis=2).sum(axis=1)
            + np.linalg.norm(drag(c, p), axis=1))


def integrate(c: Cluster, forces: np.ndarray, p: Params, dt: float) -> None:
    """Semi-implicit Euler with damping and the 0.1 m/s speed clamp
    (foam_dynamics.rs:168-196)."""
    mass = np.maximum(film_mass(c, p), 1e-9)
    v = c.vel + forces / mass[:, None] * dt
    v = v * p.damping
    speed = np.linalg.norm(v, axis=1)
    v = np.where((speed > MAX_SPEED)[:, None], v * (MAX_SPEED / np.where(speed > 0, speed, 1.0))[:, None], v)
    c.vel = v
    c.pos = c.pos + v * dt


def update_connections(c: Cluster, surface_tension: float) -> None:
    """update_connections (foam.rs:321-360) with BubbleConnection::new
    (foam.rs:106-139): pairs i < j in vector order whose overlap exceeds
    -0.3 r_i; neighbour lists in connection order."""
    n = len(c)
    conns = []
    for i in range(n):
        d = c.pos[i][None, :] - c.pos[i + 1:]
        dist = np.sqrt(np.sum(d * d, axis=1))
        overlap = c.radius[i] + c.radius[i + 1:] - dist
        for k in np.nonzero(overlap > -c.radius[i] * CONTACT_GAP)[0]:
            j = i + 1 + int(k)
            delta = c.pos[j] - c.pos[i]
            distance = math.sqrt(float(np.dot(delta, delta)))
            normal = delta / distance
            t = c.radius[i] / (c.radius[i] + c.radius[j])
            point = c.pos[i] + delta * t
            dp = 4.0 * surface_tension / c.radius[i] - 4.0 * surface_tension / c.radius[j]
            wall_r = 2.0 * surface_tension / dp if abs(dp) > 1e-6 else math.inf
            wall_t = (c.thickness_nm[i] + c.thickness_nm[j]) / 2.0 * 0.8
            conns.append((int(c.ids[i]), int(c.ids[j]), point, normal, wall_r, wall_t))
    c.connections = conns
    index = {int(b): k for k, b in enumerate(c.ids)}
    nb = [[] for _ in range(n)]
    for a, b, *_ in conns:
        nb[index[a]].append(b)
        nb[index[b]].append(a)
    c.neighbors = nb


def remove_bubble(c: Cluster, bid: int) -> bool:
    """remove_bubble (foam.rs:249-259): order kept, connections of the
    bubble dropped; other neighbour lists are left as they are."""
    hit = np.nonzero(c.ids == bid)[0]
    if len(hit) == 0:
        return False
    k = int(hit[0])
    if len(c.neighbors) == len(c):
        del c.neighbors[k]
    keep = np.arange(len(c)) != k
    c.ids, c.pos, c.radius, c.vel = c.ids[keep], c.pos[keep], c.radius[keep], c.vel[keep]
    c.thickness_nm, c.aspect, c.refr = c.thickness_nm[keep], c.aspect[keep], c.refr[keep]
    c.connections = [x for x in c.connections if x[0] != bid and x[1] != bid]
    return True


def add_bubble(c: Cluster, position, radius) -> int:
    """add_bubble (foam.rs:235-244): new id, appended, default properties
    (Bubble::new, foam.rs:52-63)."""
    bid = c.next_id
    c.next_id += 1
    c.ids = np.append(c.ids, bid)
    c.pos = np.vstack([c.pos, np.asarray(position, np.float64)[None, :]])
    c.radius = np.append(c.radius, radius)
    c.vel = np.vstack([c.vel, np.zeros((1, 3))])
    c.thickness_nm = np.append(c.thickness_nm, 500.0)
    c.aspect = np.append(c.aspect, 1.0)
    c.refr = np.append(c.refr, 1.33)
    if len(c.neighbors) == len(c) - 1:
        c.neighbors.append([])
    return bid


def merge_bubbles(c: Cluster, id_a: int, id_b: int):
    """merge_bubbles (foam.rs:365-399): volume-conserving radius, volume-
    weighted position and velocity, area-weighted thickness."""
    ia = int(np.nonzero(c.ids == id_a)[0][0])
    ib = int(np.nonzero(c.ids == id_b)[0][0])
    va, vb = float(volume(c.radius[ia])), float(volume(c.radius[ib]))
    vt = va + vb
    r = (vt * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    pos = (c.pos[ia] * va + c.pos[ib] * vb) / vt
    vel = (c.vel[ia] * va + c.vel[ib] * vb) / vt
    aa, ab = 4.0 * math.pi * c.radius[ia] ** 2, 4.0 * math.pi * c.radius[ib] ** 2
    th = (c.thickness_nm[ia] * aa + c.thickness_nm[ib] * ab) / (aa + ab)
    remove_bubble(c, id_a)
    remove_bubble(c, id_b)
    nid = add_bubble(c, pos, r)
    c.vel[-1] = vel
    c.thickness_nm[-1] = th
    return nid


def coalescence_pair(c: Cluster, p: Params):
    """First connection whose Weber number exceeds the threshold
    (foam_dynamics.rs:205-228), or None."""
    index = {int(b): k for k, b in enumerate(c.ids)}
    for a, b, _, normal, _, _ in c.connections:
        ia, ib = index[a], index[b]
        rel = float(np.dot(c.vel[ib] - c.vel[ia], normal))
        weber = p.density * rel * rel * c.radius[ia] / p.surface_tension
        if weber > p.coalescence_threshold:
            return a, b
    return None


def step(c: Cluster, p: Params, dt: float) -> None:
    """FoamSimulator::step (foam_dynamics.rs:166-202)."""
    f = compute_forces(c, p)
    integrate(c, f, p, dt)
    update_connections(c, p.surface_tension)
    if p.coalescence_enabled:
        pair = coalescence_pair(c, p)
        if pair is not None:
            merge_bubbles(c, *pair)

