"""GPU parity of the drainage step against the FP64 oracle.

Every comparison starts the HIP path and the oracle from bit-identical fields
(float-float exact FP64 values, ``workloads.ff_exact``).  Bars:
  * rupture mask, TFE membership, pole/phi indexing, query addressing: exact;
  * thickness h: relative L2 <= 1e-6, and the increment
    h - h0 is reported and held to a tighter relative L2 where stated;
  * Γ: FP32 state like the reference's WGSL kernel, |ΔΓ| <= 1e-6 per step.
"""
import math

import numpy as np
import pytest

from soap_bubble_simulation_amd import _lib as L
from soap_bubble_simulation_amd import workloads as W
from soap_bubble_simulation_amd.config import SimulationConfig
from soap_bubble_simulation_amd.drainage import (DrainageContext, DrainageSimulator,
                                                 GPUDrainageSimulator)

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def to_orc_params(O, p):
    q = O.Params()
    for name, _ in O.Params._fields_:
        setattr(q, name, getattr(p, name))
    return q


def cpu_params(flags=None, **kw):
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    if flags is not None:
        p.flags = flags
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def pair(O, p, nt, npf, h, g=None):
    ctx = DrainageContext(p, nt, npf, 500e-9)
    orc = O.Sim(to_orc_params(O, p), nt, npf)
    gg = np.full((nt, npf), 0.5) if g is None else g
    ctx.set_fields(h, gg)
    orc.set_fields(h, gg.astype(np.float32).astype(np.float64))
    return ctx, orc


def test_single_step_parity_default_grid(oracle):
    nt, npf = 128, 256
    h0 = W.noise_field(nt, npf, seed=1)
    p = cpu_params(flags=0)
    ctx, orc = pair(oracle, p, nt, npf, h0)
    dt = orc.cfl_dt()
    ctx.step(1, dt)
    orc.step(dt)
    hg, _ = ctx.fields()
    ho, _ = orc.fields()
    assert rel_l2(hg, ho) <= 1e-6
    assert rel_l2(hg - h0, ho - h0) <= 1e-5, rel_l2(hg - h0, ho - h0)


@pytest.mark.parametrize("nt,npf", [(64, 128), (40, 96), (37, 74), (9, 6), (3, 4), (20, 1030),
                                     (130, 2052)])
def test_grid_shapes_parity(oracle, nt, npf):
    """Scalar (n_phi % 4 != 0 and small grids) and float4 kernels, partial
    waves and tiles, tiny grids with both rings in one row band."""
    _shape_parity(oracle, nt, npf)


@pytest.mark.parametrize("stencil,rows", [("scalar", 1), ("scalar", 3), ("scalar", 16), ("vec4", 4), ("vec4", 8),
                                          ("vec4", 16)])
@pytest.mark.parametrize("nt,npf", [(40, 96), (130, 2052)])
def test_stencil_kernels_parity(oracle, monkeypatch, stencil, rows, nt, npf):
    """Every stencil kernel (scalar with 1-16 rows per block, float4 with 4-
    and 8-row blocks, 16-row bands), forced through the selection overrides,
    on a small and a multi-tile grid."""
    monkeypatch.setenv("SOAPFILM_STENCIL", stencil)
    monkeypatch.setenv("SOAPFILM_ROWS_PER_BLOCK", str(rows))
    _shape_parity(oracle, nt, npf)


def _shape_parity(oracle, nt, npf):
    h0 = W.noise_field(nt, npf, seed=nt * 1000 + npf, rel=0.05)
    p = cpu_params(flags=L.FLAG_SURFACTANT | L.FLAG_DIAG)
    g0 = W.ff_exact(np.clip(0.5 + 0.05 * np.random.default_rng(1).standard_normal((nt, npf)), 0, 1))
    g0 = g0.astype(np.float32).astype(np.float64)
    ctx, orc = pair(oracle, p, nt, npf, h0, g0)
    dt = 0.5 * orc.cfl_dt()
    for k in range(5):
        ctx.step(1, dt)
        orc.step(dt)
    hg, gg = ctx.fields()
    ho, go = orc.fields()
    assert rel_l2(hg, ho) <= 1e-6
    assert rel_l2(hg - h0, ho - h0) <= 1e-5
    assert np.max(np.abs(gg - go)) <= 1e-5
    # The fused FP64 film-volume reduction equals the oracle's FP64 volume of
    # the GPU's own field (same data, different summation order).
    m, _ = ctx.history(5)
    same = oracle.Sim(to_orc_params(oracle, p), nt, npf)
    same.set_fields(hg)
    assert abs(m[-1] - same.mass()) <= 1e-12 * same.mass()


def test_rupture_mask_bit_exact(oracle):
    nt, npf = 64, 128
    rng = np.random.default_rng(3)
    h0 = W.noise_field(nt, npf, seed=2)
    p = cpu_params(flags=0)
    hc = p.critical_thickness
    # Cells below, at and just above the critical thickness.
    idx = rng.choice(nt * npf, 400, replace=False)
    vals = hc * np.array([0.5, 0.999999, 1.0, 1.0000001, 0.0])
    h0.flat[idx] = W.ff_exact(np.resize(vals, idx.size))
    h0.flat[idx[:7]] = W.ff_exact(hc * (1 + np.array([-1e-12, 1e-12, -3e-15, 3e-15, 0, 2e-9, -2e-9])))
    ctx, orc = pair(oracle, p, nt, npf, h0)
    np.testing.assert_array_equal(ctx.rupture_mask(), orc.rupture_mask())
    ctx.step(1, 0.05)
    orc.step(0.05)
    hg, _ = ctx.fields()
    ho, _ = orc.fields()
    frozen = (h0 < hc)
    frozen[0, :] = False
    frozen[-1, :] = False
    np.testing.assert_array_equal(hg[frozen], h0[frozen])
    np.testing.assert_array_equal(ctx.rupture_mask(), orc.rupture_mask())
    assert rel_l2(hg, ho) <= 1e-6
    assert (hg >= 0).all()


def test_marangoni_surfactant_parity(oracle):
    nt, npf = 96, 192
    h0 = W.noise_field(nt, npf, seed=5, rel=0.05)
    rng = np.random.default_rng(6)
    g0 = np.clip(0.5 + 0.2 * W.sh_noise(nt, npf, 6, seed=9) + 0.05 * rng.standard_normal((nt, npf)),
                 0, 1).astype(np.float32).astype(np.float64)
    p = L.Params()
    L.check(L.lib().sf_params_gpu_default(p))
    p.flags = L.FLAG_SURFACTANT | L.FLAG_MARANGONI | L.FLAG_POLES_COPY_H
    p.marangoni_coeff = 1e4          # make the Marangoni term visible in h
    ctx, orc = pair(oracle, p, nt, npf, h0, g0)
    dt = 0.5 * orc.cfl_dt()
    for _ in range(3):
        ctx.step(1, dt)
        orc.step(dt)
    hg, gg = ctx.fields()
    ho, go = orc.fields()
    assert rel_l2(hg, ho) <= 1e-6
    assert rel_l2(hg - h0, ho - h0) <= 1e-5
    print("marangoni: rel_l2(h)=%.3g rel_l2(dh)=%.3g max|dG|=%.3g max|G-G0|=%.3g" % (
        rel_l2(hg, ho), rel_l2(hg - h0, ho - h0), np.max(np.abs(gg - go)), np.max(np.abs(gg - g0))))
    assert np.max(np.abs(gg - go)) <= 3e-6
    assert np.max(np.abs(gg - g0)) > 1e-3      # Γ actually evolved


def test_pole_ring_mean(oracle):
    nt, npf = 40, 96
    h0 = W.noise_field(nt, npf, seed=11, rel=0.1)
    p = cpu_params(flags=0)
    ctx, orc = pair(oracle, p, nt, npf, h0)
    dt = orc.cfl_dt()
    for k in range(4):
        ctx.step(1, dt)
        orc.step(dt)
        hg, _ = ctx.fields()
        ho, _ = orc.fields()
        assert np.all(hg[0] == hg[0, 0]) and np.all(hg[-1] == hg[-1, 0])
        assert abs(hg[0, 0] - hg[1].mean()) <= 1e-13 * hg[0, 0]
        assert abs(hg[-1, 0] - hg[-2].mean()) <= 1e-13 * hg[-1, 0]
        # The pole is the ring mean, so it can differ from the oracle's by no
        # more than the mean ring difference (FP32 increments) plus rounding.
        for pole, ring in ((0, 1), (-1, -2)):
            bound = np.mean(np.abs(hg[ring] - ho[ring])) + 1e-14 * ho[pole, 0]
            assert abs(hg[pole, 0] - ho[pole, 0]) <= bound


def test_trajectory_with_fronts(oracle):
    """128 x 256, CFL dt, TFE on, 300 steps: fronts spawn from t = 0.5 s."""
    nt, npf = 128, 256
    h0 = W.noise_field(nt, npf, seed=21)
    p = cpu_params()                         # TFE on, poles averaged
    ctx, orc = pair(oracle, p, nt, npf, h0)
    dt = orc.cfl_dt()
    steps = 300
    ctx.step(steps, dt)
    orc.step(dt, steps)
    assert ctx.current_time() == orc.current_time()
    assert ctx.active_tfe_count() == orc.active_tfe_count() > 0
    np.testing.assert_array_equal(ctx.tfe_mask(), orc.tfe_mask())
    hg, _ = ctx.fields()
    ho, _ = orc.fields()
    assert rel_l2(hg, ho) <= 1e-6, rel_l2(hg, ho)
    assert rel_l2(hg - h0, ho - h0) <= 1e-5, rel_l2(hg - h0, ho - h0)


def test_query_bit_exact(oracle):
    nt, npf = 64, 128
    h0 = W.noise_field(nt, npf, seed=31, rel=0.2)
    p = cpu_params(flags=0)
    ctx, orc = pair(oracle, p, nt, npf, h0)
    rng = np.random.default_rng(0)
    th = np.concatenate([rng.uniform(-0.5, 3.7, 500), [0.0, math.pi, math.pi / 2]])
    ph = np.concatenate([rng.uniform(-20, 20, 500), [0.0, 2 * math.pi, -1e-300]])
    got = ctx.get_thickness_batch(th, ph)
    want = np.array([orc.get_thickness(a, b) for a, b in zip(th, ph)])
    np.testing.assert_array_equal(got, want)


def test_stats_and_history(oracle):
    nt, npf = 64, 128
    h0 = W.noise_field(nt, npf, seed=41, rel=0.05)
    p = cpu_params(flags=L.FLAG_DIAG)
    ctx, orc = pair(oracle, p, nt, npf, h0)
    s = ctx.stats()
    assert s.min_thickness == orc.min_thickness() and s.max_thickness == orc.max_thickness()
    assert abs(s.mass - orc.mass()) <= 1e-12 * orc.mass()
    assert abs(s.dt_cfl - orc.cfl_dt()) <= 1e-12 * orc.cfl_dt()
    dt = orc.cfl_dt()
    masses, cfls = [], []
    for _ in range(5):
        ctx.step(1, dt)
        orc.step(dt)
        masses.append(orc.mass())
        cfls.append(orc.cfl_dt())
    m, c = ctx.history(5)
    np.testing.assert_allclose(m, masses, rtol=1e-9)
    np.testing.assert_allclose(c, cfls, rtol=1e-9)


def test_uniform_field_closed_form():
    """No diffusion, TFE off: every row follows forward Euler on
    h' = -c h^3 sin(theta) and stays uniform along phi; the trajectory tracks
    the exact solution h0 / sqrt(1 + 2 c h0^2 sin(theta) t)."""
    nt, npf = 33, 64
    p = cpu_params(flags=0, diffusion_coeff=0.0)
    ctx = DrainageContext(p, nt, npf, 500e-9)
    dt, steps = 2000.0, 20
    ctx.step(steps, dt)
    h, _ = ctx.fields()
    assert np.all(h == h[:, :1])
    c = p.density * p.gravity / (3 * p.viscosity)
    theta = np.arange(nt) * math.pi / (nt - 1)
    hh = np.full(nt, 500e-9)
    for _ in range(steps):
        hh = hh - dt * c * hh ** 3 * np.sin(theta)
        hh[0] = hh[1]
        hh[-1] = hh[-2]
    np.testing.assert_allclose(h[:, 0], hh, rtol=1e-6)
    exact = 500e-9 / np.sqrt(1 + 2 * c * (500e-9) ** 2 * np.sin(theta) * steps * dt)
    np.testing.assert_allclose(h[1:-1, 0], exact[1:-1], rtol=1e-3)
    assert h[16, 0] < 0.98 * 500e-9


def test_phi_roll_equivariance():
    nt, npf = 48, 160
    h0 = W.noise_field(nt, npf, seed=51, rel=0.1)
    p = cpu_params(flags=0)
    a = DrainageContext(p, nt, npf, 500e-9)
    b = DrainageContext(p, nt, npf, 500e-9)
    a.set_fields(h0, np.full_like(h0, 0.5))
    b.set_fields(np.roll(h0, 37, axis=1), np.full_like(h0, 0.5))
    a.step(1, 0.05)
    b.step(1, 0.05)
    ha, _ = a.fields()
    hb, _ = b.fields()
    np.testing.assert_array_equal(np.roll(ha, 37, axis=1)[1:-1], hb[1:-1])
    np.testing.assert_allclose(np.roll(ha, 37, axis=1)[[0, -1]], hb[[0, -1]], rtol=1e-14)


def test_reference_unit_tests():
    """drainage.rs:636-719 against the GPU-backed simulator."""
    config = SimulationConfig()
    sim = DrainageSimulator(config)
    assert abs(sim.current_time() - 0.0) < np.finfo(float).eps
    assert not sim.has_critical_region()
    t1 = sim.get_thickness(0.5, 1.0)
    t2 = sim.get_thickness(1.5, 3.0)
    assert abs(t1 - t2) < 1e-15
    assert abs(sim.get_thickness(1.0, 0.0) - sim.get_thickness(1.0, 2 * math.pi)) < 1e-12
    top0 = sim.get_thickness(0.1, 0.0)
    sim.step(0.001)
    assert abs(sim.current_time() - 0.001) < np.finfo(float).eps
    for _ in range(99):
        sim.step(config.dt)
    assert sim.get_thickness(0.1, 0.0) < top0
    for _ in range(10):
        sim.step(config.dt)
    sim.reset(600e-9)
    assert abs(sim.current_time()) < np.finfo(float).eps
    assert abs(sim.get_thickness(0.5, 0.5) - 600e-9) < 1e-15


def test_gpu_simulator_defaults():
    g = GPUDrainageSimulator(500e-9, 128, 64)
    assert g.grid_size() == (128, 64)
    g.step(1 / 60)
    assert g.current_time() == 0.0          # disabled by default
    g.enabled = True
    g.step(1 / 60)
    assert g.current_time() > 0.0
    h = g.current_thickness()
    assert h.shape == (64, 128) and np.isfinite(h).all()
    g.reset(400e-9)
    assert np.all(g.current_thickness() == np.float32(400e-9))


def test_large_grid_uniform_properties():
    """2048 x 4096: uniform start stays phi-uniform; FP64 mass equals 4 pi R^2 h."""
    nt, npf = 2048, 4096
    p = cpu_params(flags=L.FLAG_DIAG | L.FLAG_SURFACTANT)
    ctx = DrainageContext(p, nt, npf, 500e-9)
    s = ctx.stats()
    R = p.bubble_radius
    assert abs(s.mass - 4 * math.pi * R * R * 500e-9) <= 1e-9 * s.mass
    dth, dph = math.pi / (nt - 1), 2 * math.pi / npf
    sin1 = math.sin(dth)
    cfl = R * R / (2 * p.diffusion_coeff * (1 / dth ** 2 + 1 / (sin1 ** 2 * dph ** 2)))
    assert abs(s.dt_cfl - cfl) <= 1e-6 * cfl
    steps = 20
    ctx.step(steps, s.dt_cfl)
    h, g = ctx.fields()
    assert np.all(h == h[:, :1])
    assert np.all(g == 0.5)
    m, _ = ctx.history(steps)
    assert m.size == steps
    # Closed-form drift of a uniform film: dM/dt = -c h^3 sum_i sin(theta_i) A_i.
    c = p.density * p.gravity / (3 * p.viscosity)
    theta = np.arange(nt) * dth
    lo_b = np.where(np.arange(nt) == 0, 0.0, theta - dth / 2)
    hi_b = np.where(np.arange(nt) == nt - 1, math.pi, theta + dth / 2)
    area = R * R * (np.cos(lo_b) - np.cos(hi_b)) * dph * npf
    drift = -c * (500e-9) ** 3 * np.sum(np.sin(theta[1:-1]) * area[1:-1]) * steps * s.dt_cfl
    assert abs((m[-1] - s.mass) - drift) <= 0.05 * abs(drift) + 1e-12 * s.mass
