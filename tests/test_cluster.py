"""Foam cluster dynamics: oracle pins (CPU) and GPU parity (``-m gpu``).

The oracle (``oracle/cluster_ref.py``, FP64) is pinned by:
  * the reference's own unit tests, re-run on the oracle
    (foam.rs:422-556, foam_dynamics.rs:257-367): volume, overlap, merge
    conserves volume, default cluster has connections, step keeps the count,
    buoyancy along gravity, repulsion separates overlapping bubbles;
  * invariants a sign, index or operand slip breaks: pair forces are
    antisymmetric, so with gravity, drag and damping off the film-mass-
    weighted momentum is conserved by a step; the total pair force of a
    cluster is zero;
  * closed forms: a lone bubble under gravity and drag follows the damped
    Euler recursion in closed form and saturates at the 0.1 m/s clamp;
    van der Waals magnitude falls as 1/d^2, Hertzian repulsion grows as
    overlap^1.5, adhesion is linear in the gap (ratios, not the formulas);
  * order semantics built by hand: a chain of bubbles connects only
    neighbours, in (i, j) order; the contact gap is 30 % of the *first*
    bubble's radius, so swapping a small and a large bubble flips a
    borderline contact; merged thickness of equal radii is the mean.
The GPU tests compare the HIP kernels with the oracle: forces within 1e-5 of
each bubble's force scale (sum of magnitudes of its terms, FP32 pair forces
summed in a different order), states after steps within FP32 rounding of the
trajectory, connection lists equal where no pair sits within FP32 rounding of
the contact threshold.
"""
import math

import numpy as np
import pytest

from oracle import cluster_ref as R
from soap_bubble_simulation_amd import workloads as W


def two(ra, rb, d, p=None):
    c = R.Cluster.from_arrays([(0.0, 0.0, 0.0), (d, 0.0, 0.0)], [ra, rb])
    return c


# --------------------------------------------------------------------------
# the reference's unit tests on the oracle

def test_reference_bubble_volume_and_overlap():
    assert abs(float(R.volume(1.0)) - 4.0 / 3.0 * math.pi) < 1e-6            # foam.rs:434-439
    a = R.Cluster.from_arrays([(0, 0, 0), (0.8, 0, 0), (1.5, 0, 0)], [0.5, 0.5, 0.5])
    ov = lambda i, j: a.radius[i] + a.radius[j] - np.linalg.norm(a.pos[i] - a.pos[j])
    assert ov(0, 1) > 0 and not ov(0, 2) > 0                                   # foam.rs:442-450


def test_reference_merge_conserves_volume():
    c = R.Cluster.from_arrays([(0, 0, 0), (0.8, 0, 0)], [0.5, 0.5])           # foam.rs:495-507
    before = float(R.volume(0.5)) * 2
    nid = R.merge_bubbles(c, 0, 1)
    assert nid == 2 and len(c) == 1
    assert abs(float(R.volume(c.radius[0])) - before) < 1e-6
    assert np.allclose(c.pos[0], [0.4, 0, 0])


def test_reference_default_cluster_connections_and_step():
    c = R.default_cluster()
    assert len(c.connections) > 0                                              # foam.rs:553-555
    p = R.Params()
    f = R.compute_forces(c, p)
    assert f.shape == (5, 3)                                                   # foam_dynamics.rs:311-320
    R.step(c, p, 0.001)
    assert len(c) == 5                                                         # foam_dynamics.rs:323-331
    b = R.buoyancy(R.default_cluster(), p)[0]                                  # foam_dynamics.rs:334-347
    g = np.array(p.gravity) / np.linalg.norm(p.gravity)
    assert np.linalg.norm(b) < 1e-6 or abs(abs(np.dot(g, b / np.linalg.norm(b))) - 1.0) < 1e-12


def test_reference_repulsion_separates():
    c = two(0.03, 0.03, 0.04)                                                  # foam_dynamics.rs:350-366
    p = R.Params(gravity=(0.0, 0.0, 0.0))
    for _ in range(100):
        R.step(c, p, 0.001)
    assert np.linalg.norm(c.pos[0] - c.pos[1]) > 0.04


# --------------------------------------------------------------------------
# invariants and closed forms

def test_pair_forces_antisymmetric_and_momentum():
    pos, r, vel, th = W.cluster_bubbles(60, 3, box=0.08)
    c = R.Cluster.from_arrays(pos, r, vel, th)
    p = R.Params(gravity=(0.0, 0.0, 0.0), viscosity=0.0, damping=1.0)
    F = R.pair_forces(c, p)
    assert np.abs(F + F.transpose(1, 0, 2)).max() == 0.0
    scale = np.abs(F).sum()
    assert np.abs(F.sum(axis=(0, 1))).max() <= 1e-12 * scale
    m = R.film_mass(c, p)
    c.vel *= 1e-3      # far below the speed clamp, and a step short enough to stay there
    p0 = (m[:, None] * c.vel).sum(axis=0)
    R.integrate(c, R.compute_forces(c, p), p, 1e-8)
    p1 = (m[:, None] * c.vel).sum(axis=0)
    assert np.abs(p1 - p0).max() <= 1e-12 * (np.abs(m[:, None] * c.vel).sum() + 1e-30)


def test_lone_bubble_closed_form():
    """v_{n+1} = (v_n + a dt - k v_n dt) d with a = net weight / m, k = drag/m;
    the fixed point and the clamp at 0.1 m/s."""
    p = R.Params(viscosity=0.0)
    c = R.Cluster.from_arrays([(0, 0, 0)], [0.005], thickness_nm=[20000.0])
    m = float(R.film_mass(c, p)[0])
    weight = m * 9.81 - float(R.volume(0.005)) * 1.2 * 9.81
    assert weight > 0
    a, d, dt = weight / m, p.damping, 1e-3
    v = 0.0
    for n in range(1, 30):
        R.step(c, p, dt)
        v = (v + a * dt) * d
        assert abs(-c.vel[0, 1] - min(v, 0.1)) <= 1e-12
        if v > 0.1:
            v = 0.1
    # heavy film reaches the clamp: speed exactly 0.1 after enough steps
    p2 = R.Params(viscosity=0.0, damping=1.0)
    for _ in range(200):
        R.step(c, p2, 1e-2)
    assert abs(np.linalg.norm(c.vel[0]) - 0.1) <= 1e-12


def test_force_law_scalings():
    p = R.Params()
    f = lambda ra, rb, d: R.pair_forces(two(ra, rb, d), p)[0, 1, 0]
    # van der Waals window (s, 1.2 s excluded to keep adhesion out): 1/d^2
    s = 0.02
    assert abs(f(0.01, 0.01, 1.5 * s) / f(0.01, 0.01, 2.5 * s) - (2.5 / 1.5) ** 2) < 1e-12
    # s < d < 1.2 s: van der Waals plus adhesion, the latter linear in 1.2 s - d
    vdw = lambda d: p.van_der_waals_strength * 0.01 * 0.01 / d ** 2
    a1, a2 = f(0.01, 0.01, 1.05 * s) - vdw(1.05 * s), f(0.01, 0.01, 1.15 * s) - vdw(1.15 * s)
    assert abs(a1 / a2 - (1.2 - 1.05) / (1.2 - 1.15)) < 1e-9
    # deep overlap (< 0.95 s): only repulsion, magnitude ~ overlap^1.5, pushes apart
    r1, r2 = f(0.01, 0.01, 0.5 * s), f(0.01, 0.01, 0.8 * s)
    assert r1 < 0 and r2 < 0 and abs(r1 / r2 - (0.5 / 0.2) ** 1.5) < 1e-12
    assert f(0.01, 0.01, 3.5 * s) == 0.0 and a1 > 0


def test_connection_order_and_asymmetric_gap():
    # chain along x: only consecutive bubbles touch
    xs = [0.0, 0.019, 0.038, 0.057]
    c = R.Cluster.from_arrays([(x, 0, 0) for x in xs], [0.01] * 4)
    R.update_connections(c, 0.025)
    assert [(a, b) for a, b, *_ in c.connections] == [(0, 1), (1, 2), (2, 3)]
    assert c.neighbors == [[1], [0, 2], [1, 3], [2]]
    # gap 5 mm: contact iff 0.3 r_first > 5 mm
    small_first = R.Cluster.from_arrays([(0, 0, 0), (0.035, 0, 0)], [0.01, 0.02])
    big_first = R.Cluster.from_arrays([(0, 0, 0), (0.035, 0, 0)], [0.02, 0.01])
    R.update_connections(small_first, 0.025)
    R.update_connections(big_first, 0.025)
    assert len(small_first.connections) == 0 and len(big_first.connections) == 1
    # contact point divides the centre line in the ratio of the radii
    _, _, point, normal, wall_r, wall_t = big_first.connections[0]
    assert np.allclose(point, [0.035 * 2 / 3, 0, 0]) and np.allclose(normal, [1, 0, 0])
    assert wall_t == pytest.approx(500.0 * 0.8) and wall_r < 0
    same = R.Cluster.from_arrays([(0, 0, 0), (0.02, 0, 0)], [0.01, 0.01])
    R.update_connections(same, 0.025)
    assert same.connections[0][4] == math.inf


def test_merge_thickness_and_coalescence_order():
    c = R.Cluster.from_arrays([(0, 0, 0), (0.015, 0, 0), (0.5, 0, 0)], [0.01, 0.01, 0.01],
                              vel=[(0.05, 0, 0), (-0.05, 0, 0), (0, 0, 0)], thickness_nm=[300.0, 700.0, 500.0])
    p = R.Params(coalescence_enabled=True, gravity=(0, 0, 0))
    R.update_connections(c, p.surface_tension)
    assert R.coalescence_pair(c, p) == (0, 1)
    R.merge_bubbles(c, 0, 1)
    assert list(c.ids) == [2, 3] and c.thickness_nm[1] == pytest.approx(500.0)
    assert c.radius[1] == pytest.approx(0.01 * 2 ** (1 / 3))


# --------------------------------------------------------------------------
# GPU parity

def _sim(pos, r, vel=None, th=None, **params):
    from soap_bubble_simulation_amd.cluster import FoamSimulator
    s = FoamSimulator(pos, r, vel, th)
    for k, v in params.items():
        setattr(s, k, v)
    return s


def _oracle_params(**kw):
    f32 = lambda x: float(np.float32(x))
    d = dict(surface_tension=f32(0.025), viscosity=f32(0.001), density=1000.0,
             gravity=(0.0, f32(-9.81), 0.0), van_der_waals_strength=f32(0.001), repulsion_stiffness=5.0,
             damping=f32(0.8), coalescence_enabled=False, coalescence_threshold=0.5)
    d.update(kw)
    return R.Params(**d)


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed,box", [(7, 1, 0.05), (300, 2, 0.12), (1500, 3, 0.25)])
def test_gpu_forces_match_oracle(n, seed, box):
    pos, r, vel, th = W.cluster_bubbles(n, seed, box=box)
    s = _sim(pos, r, vel, th)
    got = s.compute_forces().astype(np.float64)
    c = R.Cluster.from_arrays(pos, r, vel, th)
    p = _oracle_params()
    want = R.compute_forces(c, p)
    scale = R.force_scale(c, p)
    err = np.abs(got - want).max(axis=1) / np.maximum(scale, 1e-30)
    print("cluster n=%d: max force error / scale = %.3g" % (n, err.max()))
    assert err.max() <= 1e-5


@pytest.mark.gpu
def test_gpu_default_cluster_matches_reference_behaviour():
    from soap_bubble_simulation_amd.cluster import FoamSimulator
    s = FoamSimulator.with_default_cluster()
    c = R.default_cluster()
    assert s.bubble_count() == 5 and s.connection_count() == len(c.connections) > 0
    assert [(x.bubble_a, x.bubble_b) for x in s.connections()] == [(a, b) for a, b, *_ in c.connections]
    assert s.neighbors() == c.neighbors
    p = _oracle_params()
    for _ in range(11):                                                       # foam.rs:510-529
        s.step(0.016)
        R.step(c, p, 0.016)
    b = s.bubbles()
    assert np.abs(b["position"] - c.pos).max() <= 1e-6
    assert np.abs(b["velocity"] - c.vel).max() <= 1e-5
    assert [(x.bubble_a, x.bubble_b) for x in s.connections()] == [(a, b) for a, b, *_ in c.connections]


@pytest.mark.gpu
def test_gpu_repulsion_separates():
    s = _sim([(0, 0, 0), (0.04, 0, 0)], [0.03, 0.03], gravity=np.zeros(3))    # foam_dynamics.rs:350-366
    s.step_many(100, 0.001)
    b = s.bubbles()
    assert np.linalg.norm(b["position"][0] - b["position"][1]) > 0.04


def _ambiguous_pairs(c, tol=1e-5):
    """Id pairs (i < j) within tol (relative) of the contact threshold, whose
    FP32 and FP64 contact decisions may differ."""
    d = np.sqrt(((c.pos[:, None, :] - c.pos[None, :, :]) ** 2).sum(axis=2))
    ov = c.radius[:, None] + c.radius[None, :] - d + c.radius[:, None] * R.CONTACT_GAP
    iu = np.triu_indices(len(c), 1)
    near = np.abs(ov[iu]) <= tol * (c.radius[:, None] + c.radius[None, :])[iu]
    return {(int(c.ids[i]), int(c.ids[j])) for i, j in zip(iu[0][near], iu[1][near])}


def _step_tolerances(c, p, dt):
    """Per-bubble bounds on |v_gpu - v| and |x_gpu - x| for one step from the
    same state: the force error (1e-5 of the force scale, FP32 pair terms
    summed in another order) through the Euler update, or through the
    direction of a clamped velocity, plus FP32 rounding."""
    F = R.compute_forces(c, p)
    eF = 1e-5 * R.force_scale(c, p) + 1e-6 * np.linalg.norm(F, axis=1)
    m = np.maximum(R.film_mass(c, p), 1e-9)
    dv = p.damping * dt * eF / m
    vu = np.linalg.norm((c.vel + F / m[:, None] * dt) * p.damping, axis=1)
    tol_v = np.where(vu > R.MAX_SPEED * (1 - 1e-6), 2.0 * R.MAX_SPEED * dv / np.maximum(vu, 1e-30), dv)
    tol_v = tol_v + 4e-7 * np.maximum(vu, 1e-3)
    tol_x = tol_v * dt + 4e-7 * np.abs(c.pos).max(axis=1) + 1e-9
    return tol_v, tol_x


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed,box", [(200, 11, 0.2), (1024, 12, 0.3)])
def test_gpu_steps_and_connections_match_oracle(n, seed, box):
    """Ten steps, each checked against the oracle stepped from the GPU's own
    previous state, then the connection list of the final state."""
    pos, r, vel, th = W.cluster_bubbles(n, seed, box=box)
    s = _sim(pos, r, vel, th)
    p = _oracle_params()
    dt = 1e-3
    worst = 0.0
    for _ in range(10):
        b = s.bubbles()
        c = R.Cluster.from_arrays(b["position"], b["radius"], b["velocity"], b["thickness_nm"], b["ids"])
        tol_v, tol_x = _step_tolerances(c, p, dt)
        R.step(c, p, dt)
        s.step(dt)
        g = s.bubbles()
        ev = np.abs(g["velocity"] - c.vel).max(axis=1)
        ex = np.abs(g["position"] - c.pos).max(axis=1)
        worst = max(worst, float((ev / tol_v).max()), float((ex / tol_x).max()))
        assert np.all(ev <= tol_v) and np.all(ex <= tol_x)
    print("cluster n=%d: worst error / bound over 10 steps = %.3g" % (n, worst))
    b = s.bubbles()
    g = R.Cluster.from_arrays(b["position"], b["radius"], b["velocity"], b["thickness_nm"], b["ids"])
    R.update_connections(g, p.surface_tension)
    amb = _ambiguous_pairs(g)
    conns = [x for x in s.connections() if (x.bubble_a, x.bubble_b) not in amb]
    want = [x for x in g.connections if (x[0], x[1]) not in amb]
    assert len(conns) > 0
    assert [(x.bubble_a, x.bubble_b) for x in conns] == [(a, bb) for a, bb, *_ in want]
    radius = dict(zip(g.ids.tolist(), g.radius.tolist()))
    for x, (a, bb, point, normal, wall_r, wall_t) in zip(conns, want):
        assert np.abs(x.contact_point - point).max() <= 1e-6
        assert np.abs(x.contact_normal - normal).max() <= 1e-5
        assert x.wall_thickness_nm == pytest.approx(wall_t, rel=1e-6)
        # 2 gamma / (4 gamma / r_a - 4 gamma / r_b): FP32 rounding of the two
        # pressures is amplified by their cancellation
        pa, pb = 1.0 / radius[a], 1.0 / radius[bb]
        rel = 1e-6 + 4e-7 * (pa + pb) / max(abs(pa - pb), 1e-30)
        assert (math.isinf(wall_r) and math.isinf(x.wall_curvature_radius)) or \
            x.wall_curvature_radius == pytest.approx(wall_r, rel=rel)
    if not amb:
        assert s.neighbors() == g.neighbors


@pytest.mark.gpu
def test_gpu_coalescence_and_cluster_edits():
    pos = [(0, 0, 0), (0.015, 0, 0), (0.5, 0, 0)]
    vel = [(0.05, 0, 0), (-0.05, 0, 0), (0, 0, 0)]
    s = _sim(pos, [0.01, 0.01, 0.01], vel, [300.0, 700.0, 500.0], coalescence_enabled=True,
             gravity=np.zeros(3))
    c = R.Cluster.from_arrays(pos, [0.01, 0.01, 0.01], vel, [300.0, 700.0, 500.0])
    p = _oracle_params(coalescence_enabled=True, gravity=(0.0, 0.0, 0.0))
    s.step(1e-3)
    R.step(c, p, 1e-3)
    b = s.bubbles()
    assert list(b["ids"]) == list(c.ids) == [2, 3]
    assert np.abs(b["position"] - c.pos).max() <= 1e-7
    assert np.abs(b["radius"] - c.radius).max() <= 1e-8
    assert np.abs(b["thickness_nm"] - c.thickness_nm).max() <= 1e-3
    # the step's connections less the merged pair, and stale neighbour lists
    assert [(x.bubble_a, x.bubble_b) for x in s.connections()] == [(a, bb) for a, bb, *_ in c.connections]
    assert s.neighbors() == c.neighbors
    nid = s.add_bubble((0.1, 0.1, 0.1), 0.02)
    assert nid == 4 and s.bubble_count() == 3 and s.remove_bubble(2) and not s.remove_bubble(2)
    assert list(s.bubbles()["ids"]) == [3, 4]


@pytest.mark.gpu
def test_gpu_cluster_feeds_foam_drainage():
    from soap_bubble_simulation_amd import _lib as L
    from soap_bubble_simulation_amd.foam import FoamDrainage
    pos, r, vel, th = W.cluster_bubbles(24, 5, box=0.2)
    s = _sim(pos, r, vel, th)
    fb = s.foam_bubbles(16, 64, 2.5e-3, 25e-3)
    assert [b.bubble_radius for b in fb] == pytest.approx(list(r.astype(np.float64)))
    assert [b.h0 for b in fb] == pytest.approx(list(th.astype(np.float64) * 1e-9), rel=1e-7)
    u = (np.log(r.astype(np.float64)) - math.log(2.5e-3)) / (math.log(25e-3) - math.log(2.5e-3))
    assert [b.n_theta for b in fb] == [int(round(16 + 48 * x)) for x in np.clip(u, 0, 1)]
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    p.flags = L.FLAG_SURFACTANT
    foam = FoamDrainage(p, fb)
    # uniform films: the foam's mean thickness is h0, so the round trip keeps it
    s.thickness_from_foam(foam)
    assert np.abs(s.bubbles()["thickness_nm"] - th).max() <= 1e-3
    foam.step(50)
    mass = np.zeros(len(fb))
    L.check(L.lib().sf_foam_stats(foam.handle, L.dptr(mass), None, None, None))
    s.thickness_from_foam(foam)
    want = mass / (4 * math.pi * r.astype(np.float64) ** 2) * 1e9
    got = s.bubbles()["thickness_nm"]
    assert np.abs(got - want).max() <= 1e-3 * want.max()
    assert np.all(got <= th * (1 + 1e-6)) and np.any(got < th)      # the films drain


@pytest.mark.gpu
def test_gpu_graph_replay_matches_single_steps():
    """40 steps as two 16-step graph replays plus 8 launches equal 40 single
    steps bit for bit (same kernels, same order)."""
    pos, r, vel, th = W.cluster_bubbles(700, 21, box=0.2)
    a = _sim(pos, r, vel, th)
    b = _sim(pos, r, vel, th)
    a.step_many(40, 1e-3)
    for _ in range(40):
        b.step(1e-3)
    ba, bb = a.bubbles(), b.bubbles()
    assert np.array_equal(ba["position"], bb["position"]) and np.array_equal(ba["velocity"], bb["velocity"])
    assert [(x.bubble_a, x.bubble_b) for x in a.connections()] == [(x.bubble_a, x.bubble_b) for x in b.connections()]


@pytest.mark.gpu
def test_gpu_cluster_edge_cases():
    from soap_bubble_simulation_amd.cluster import FoamSimulator
    empty = FoamSimulator()
    empty.step_many(20, 1e-3)
    assert empty.bubble_count() == 0 and empty.connection_count() == 0
    assert empty.compute_forces().shape == (0, 3) and empty.neighbors() == []
    # a lone bubble: buoyancy and drag only, as the oracle
    one = _sim([(0.01, 0.02, 0.03)], [0.004], [(0.01, 0.0, 0.0)], [20000.0])
    c = R.Cluster.from_arrays([(0.01, 0.02, 0.03)], [0.004], [(0.01, 0.0, 0.0)], [20000.0])
    p = _oracle_params()
    for _ in range(5):
        one.step(1e-3)
        R.step(c, p, 1e-3)
    b = one.bubbles()
    assert np.abs(b["velocity"] - c.vel).max() <= 1e-6 and one.connection_count() == 0
    # coincident centres: no interaction force (distance < 1e-6, foam_dynamics.rs:120-122)
    two = _sim([(0, 0, 0), (0, 0, 0)], [0.01, 0.01], gravity=np.zeros(3))
    assert np.abs(two.compute_forces()).max() == 0.0
    assert two.connection_count() == 1
