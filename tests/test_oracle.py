"""Pins of the FP64 CPU oracle (CPU only).

The reference is a Rust crate that cannot be built here, and it ships no
numeric fixtures for this path, so every oracle function is pinned by the
reference's own unit-test assertions, by closed forms or by an independent
formulation:

drainage (oracle/sf_oracle_drainage.c)
  * reference asserts drainage.rs:621-719 .......... test_reference_drainage_asserts
  * gravity sink: exact ODE solution of h' = -c h^3 sin(theta) ... test_sink_ode
  * phi stencil: discrete eigenvalue of cos(m phi) ... test_phi_eigenvalue
  * theta stencil + cot term + 1/R^2: Y_l^m eigenvalues -l(l+1)/R^2 ... test_sphere_laplacian
  * rupture freeze, h >= 0 clamp .................... test_rupture_and_clamp
  * pole ring mean (exact sum) ...................... test_pole_mean
  * TFE schedule, ranges, separability, factor ...... test_tfe_schedule
  * bilinear query exact on linear fields, periodic . test_query_linear
  * mass: sum of areas = 4 pi R^2 .................. test_mass_closed_form
  * CFL: 0.1151 s at 128x256; growth above the bound test_cfl
interference (oracle/sf_oracle_interference.c)
  * reference asserts interference.rs:619-780 ...... test_reference_interference_asserts
  * Fresnel vs the sin/tan textbook forms .......... test_fresnel_textbook
  * Airy special phases 2 pi m and (2m+1) pi ........ test_airy_special_phases
  * reference asserts interference_lut.rs:170-204 .. test_reference_lut_asserts
  * runtime LUT reflectance stage (interference_lut.rs:45-133):
      t = 0 column: phase pi, a = 4R/(1+R)^2 with R from the sin/tan
      Fresnel forms; column ratios across angles .... test_lut_zero_thickness_column
      cos = 0 row: R = 1, F = 4e6, a = F s^2/(1+F s^2) test_lut_grazing_row
      per-wavelength Airy term = the phase-pinned
      FP64 calculate_reflected_intensity (cos-delta
      form, interference.rs:312-353) pointwise ...... test_lut_airy_matches_reflected_intensity
  * runtime LUT colour stage: linear in the Airy
    terms; curves recovered through the published
    sRGB -> XYZ matrix are the CIE 1931 observer's  test_lut_colour_stage
  * spectral RGB: a one-line colour curve gives the
    phase-pinned FP64 reflected intensity / 81 ...... test_spectral_one_line_curves
  * spectral RGB reduces to the runtime LUT for a
    7-line colour curve (two independent paths) ..... test_spectral_reduces_to_lut
  * view angle geometry ............................ test_view_cos
test_runtime_lut_numpy compares the runtime LUT with a numpy restatement of
the same formulas; it guards against slips in the C transcription only and
pins nothing by itself.  The coefficients of the reference's Gaussian fits
to the colour-matching functions (interference_lut.rs:136-148) are data of
the reference that no published table reproduces exactly (they take one
width per lobe of the two-sided fits of Wyman, Sloan and Shirley 2013): they
are pinned in shape only (test_lut_colour_stage), and a slip in a small
coefficient is caught by that transcription cross-check alone.
Marangoni and surfactant terms (drainage.wgsl:150-201) are pinned by
test_surfactant_modes: a Γ = cos(m phi) row decays with the discrete
eigenvalue, and the Marangoni term equals m gamma_r |grad Γ| h^2 for a
linear-in-phi Γ.
"""
import math

import numpy as np
import pytest

EPS = np.finfo(float).eps


def make(O, nt=128, npf=256, flags=None, **kw):
    p = O.cpu_default_params()
    if flags is not None:
        p.flags = flags
    for k, v in kw.items():
        setattr(p, k, v)
    return O.Sim(p, nt, npf), p


# ---------------------------------------------------------------- drainage

def test_reference_drainage_asserts(oracle):
    s, p = make(oracle, 64, 128)
    h, _ = s.fields()
    assert h.shape == (64, 128) and abs(h[0, 0] - 500e-9) < EPS          # 621-626
    s2, _ = make(oracle)
    assert abs(s2.current_time()) < EPS and not s2.has_critical_region()  # 636-642
    s2.step(0.001)
    assert abs(s2.current_time() - 0.001) < EPS                          # 645-653
    s3, _ = make(oracle)
    assert abs(s3.get_thickness(0.5, 1.0) - s3.get_thickness(1.5, 3.0)) < 1e-15   # 656-665
    top = s3.get_thickness(0.1, 0.0)
    for _ in range(100):
        s3.step(0.001)
    assert s3.get_thickness(0.1, 0.0) < top                               # 668-688
    assert abs(s3.get_thickness(1.0, 0.0) - s3.get_thickness(1.0, 2 * math.pi)) < 1e-12  # 691-702
    s3.reset(600e-9)
    assert abs(s3.current_time()) < EPS                                   # 705-719
    assert abs(s3.get_thickness(0.5, 0.5) - 600e-9) < 1e-15


def test_sink_ode(oracle):
    nt, npf = 33, 16
    s, p = make(oracle, nt, npf, flags=0, diffusion_coeff=0.0)
    c = p.density * p.gravity / (3 * p.viscosity)
    dt, steps = 200.0, 200
    s.step(dt, steps)
    h, _ = s.fields()
    assert np.all(h == h[:, :1])
    theta = np.arange(nt) * math.pi / (nt - 1)
    exact = 500e-9 / np.sqrt(1 + 2 * c * (500e-9) ** 2 * np.sin(theta) * dt * steps)
    # forward Euler global error is O(dt): relative 1e-4 here
    np.testing.assert_allclose(h[1:-1, 0], exact[1:-1], rtol=2e-4)
    assert h[16, 0] < 0.97 * 500e-9


def test_phi_eigenvalue(oracle):
    nt, npf, m = 41, 64, 5
    s, p = make(oracle, nt, npf, flags=0, gravity=0.0)
    phi = np.arange(npf) * 2 * math.pi / npf
    eps = 1e-9
    h0 = 500e-9 + eps * np.tile(np.cos(m * phi), (nt, 1))
    s.set_fields(h0)
    dt = 1e-3
    s.step(dt)
    h, _ = s.fields()
    dth, dph = math.pi / (nt - 1), 2 * math.pi / npf
    R2 = p.bubble_radius ** 2
    for i in (1, 5, 20, 39):
        sin_t = math.sin(i * dth)
        lam = (2 - 2 * math.cos(m * dph)) / (dph ** 2 * sin_t ** 2) / R2
        amp = np.dot(h[i] - 500e-9, np.cos(m * phi)) / np.dot(np.cos(m * phi), np.cos(m * phi))
        assert abs(amp / eps - (1 - dt * p.diffusion_coeff * lam)) <= 1e-6


def test_sphere_laplacian(oracle):
    nt, npf = 257, 512
    s, p = make(oracle, nt, npf, flags=0, gravity=0.0, diffusion_coeff=1.0)
    theta = np.arange(nt)[:, None] * math.pi / (nt - 1)
    phi = np.arange(npf)[None, :] * 2 * math.pi / npf
    R2 = p.bubble_radius ** 2
    modes = {
        (1, 0): np.cos(theta) + 0 * phi,
        (2, 0): 0.5 * (3 * np.cos(theta) ** 2 - 1) + 0 * phi,
        (2, 1): np.sin(theta) * np.cos(theta) * np.cos(phi),
        (3, 2): np.sin(theta) ** 2 * np.cos(theta) * np.cos(2 * phi),
    }
    dt = 1e-12
    for (l, m), Y in modes.items():
        h0 = 1.0 + 1e-3 * Y
        s.set_fields(h0)
        s.step(dt)
        h, _ = s.fields()
        lap = (h - h0) / dt
        band = slice(nt // 8, nt - nt // 8)
        expect = -l * (l + 1) / R2 * 1e-3 * Y
        err = np.max(np.abs(lap[band] - expect[band])) / np.max(np.abs(expect[band]))
        assert err < 5e-3, ((l, m), err)
        s.reset(500e-9)


def test_rupture_and_clamp(oracle):
    s, p = make(oracle, 32, 64, flags=0, diffusion_coeff=0.0)
    rng = np.random.default_rng(0)
    h0 = 500e-9 * np.abs(1 + 0.3 * rng.standard_normal((32, 64)))
    h0[rng.random((32, 64)) < 0.2] = 5e-9
    s.set_fields(h0)
    rm = s.rupture_mask()
    np.testing.assert_array_equal(rm.astype(bool), h0 < p.critical_thickness)
    s.step(1e7)                    # huge step: the h^3 sink overshoots below zero
    h, _ = s.fields()
    inner = rm[1:-1].astype(bool)
    np.testing.assert_array_equal(h[1:-1][inner], h0[1:-1][inner])
    assert (h >= 0).all() and (h[1:-1][~inner] == 0).any()


def test_pole_mean(oracle):
    s, _ = make(oracle, 24, 80, flags=0)
    rng = np.random.default_rng(1)
    s.set_fields(500e-9 * (1 + 0.05 * rng.standard_normal((24, 80))))
    s.step(0.01)
    h, _ = s.fields()
    for pole, ring in ((0, 1), (-1, -2)):
        exact = math.fsum(h[ring]) / 80
        assert np.all(h[pole] == h[pole, 0])
        assert abs(h[pole, 0] - exact) <= 80 * EPS * exact


def test_tfe_schedule(oracle):
    nt, npf = 128, 256
    s, p = make(oracle, nt, npf)
    off, _ = make(oracle, nt, npf, flags=0)
    dt = 0.1
    seen_first = None
    for k in range(200):
        h_before, _ = s.fields()
        off.set_fields(h_before)
        s.step(dt)
        off.step(dt)
        fr = s.fronts()
        assert len(fr) <= 8
        if fr and seen_first is None:
            seen_first = k
        for phic, ext, w, v in fr:
            assert 0 <= phic <= 2 * math.pi and 0.3 <= w <= 0.7 and 0.05 <= v <= 0.1
            assert 0 < ext < 0.95 * math.pi / 2
        mask = s.tfe_mask().astype(bool)
        # separable: each row's mask is a union of the fronts' phi runs
        hon, _ = s.fields()
        hoff, _ = off.fields()
        ratio = hon[mask] / hoff[mask]
        f = 1 - (1 - 0.85) * 0.5 * dt
        np.testing.assert_allclose(ratio, f, rtol=4 * EPS)
        np.testing.assert_array_equal(hon[~mask], hoff[~mask])
    # first spawn at the first step where the FP64 countdown 0.5 - dt - dt ...
    # reaches <= 0 (0.5 - 5 * 0.1 leaves 2.8e-17 in FP64, so the sixth step)
    t, k = 0.5, 0
    while True:
        t -= dt
        if t <= 0.0:
            break
        k += 1
    assert seen_first == k == 5
    assert s.active_tfe_count() > 0


def test_query_linear(oracle):
    nt, npf = 30, 40
    s, _ = make(oracle, nt, npf, flags=0)
    ii, jj = np.meshgrid(np.arange(nt), np.arange(npf), indexing="ij")
    a, b, c = 5e-7, 3e-9, 1e-9
    s.set_fields(a + b * ii + c * jj)
    dth, dph = math.pi / (nt - 1), 2 * math.pi / npf
    rng = np.random.default_rng(2)
    for _ in range(200):
        th = rng.uniform(0, math.pi)
        ph = rng.uniform(0, 2 * math.pi - 1.5 * dph)
        want = a + b * th / dth + c * ph / dph
        assert abs(s.get_thickness(th, ph) - want) <= 1e-14 * want
    assert s.get_thickness(-1.0, 0.3) == s.get_thickness(0.0, 0.3)
    assert abs(s.get_thickness(1.2, 0.3 + 4 * math.pi) - s.get_thickness(1.2, 0.3)) <= 1e-20


def test_mass_closed_form(oracle):
    for nt, npf in ((3, 4), (64, 128), (129, 256)):
        s, p = make(oracle, nt, npf, flags=0)
        area = sum(oracle.cell_area(nt, npf, p.bubble_radius, i) for i in range(nt)) * npf
        assert abs(area - 4 * math.pi * p.bubble_radius ** 2) <= 1e-12 * area
        assert abs(s.mass() - area * 500e-9) <= 1e-12 * s.mass()


def test_cfl(oracle):
    s, p = make(oracle, 128, 256, flags=0)
    assert abs(s.cfl_dt() - 0.1151) <= 1e-4          # 128 x 256, D = 1e-9, R = 25 mm
    nt, npf = 32, 64
    rng = np.random.default_rng(3)
    noise = 500e-9 * (1 + 1e-3 * rng.standard_normal((nt, npf)))
    amp = {}
    for safety in (0.9, 1.3):
        t, _ = make(oracle, nt, npf, flags=0, gravity=0.0)
        t.set_fields(noise)
        dt = safety * t.cfl_dt()
        t.step(dt, 200)
        h, _ = t.fields()
        amp[safety] = np.max(np.abs(h - 500e-9))
    a0 = np.max(np.abs(noise - 500e-9))
    assert amp[0.9] < a0 and amp[1.3] > 100 * a0     # decays below, grows above


def test_surfactant_modes(oracle):
    nt, npf, m = 33, 64, 3
    flags = oracle.SURFACTANT | oracle.POLES_COPY_H
    s, p = make(oracle, nt, npf, flags=flags, gravity=0.0, diffusion_coeff=0.0,
                surfactant_diffusion=1e-9)
    phi = np.arange(npf) * 2 * math.pi / npf
    g0 = 0.5 + 0.1 * np.tile(np.cos(m * phi), (nt, 1))
    s.set_fields(np.full((nt, npf), 500e-9), g0)
    dt = 1.0
    s.step(dt)
    _, g = s.fields()
    dth, dph = math.pi / (nt - 1), 2 * math.pi / npf
    for i in (3, 16):
        lam = (2 - 2 * math.cos(m * dph)) / (dph ** 2 * math.sin(i * dth) ** 2) / p.bubble_radius ** 2
        amp = np.dot(g[i] - 0.5, np.cos(m * phi)) / np.dot(np.cos(m * phi), np.cos(m * phi))
        assert abs(amp / 0.1 - (1 - dt * 1e-9 * lam)) <= 1e-9
    # Marangoni: Γ linear in theta on interior rows gives |grad Γ| = k / R.
    s2, p2 = make(oracle, nt, npf, flags=oracle.MARANGONI | oracle.POLES_COPY_H, gravity=0.0,
                  diffusion_coeff=0.0, marangoni_coeff=1e3)
    theta = np.arange(nt)[:, None] * dth + 0 * phi[None, :]
    k = 0.05
    s2.set_fields(np.full((nt, npf), 500e-9), 0.2 + k * theta)
    s2.step(1.0)
    h, _ = s2.fields()
    want = 500e-9 + 1.0 * 1e3 * p2.gamma_reduction * (k / p2.bubble_radius) * (500e-9) ** 2
    np.testing.assert_allclose(h[2:-2], want, rtol=1e-9)


# ------------------------------------------------------------ interference

def test_reference_interference_asserts(oracle):
    O = oracle
    assert abs(O.transmission_cos(1.33, 1.0) - 1.0) < 1e-6                     # 649-654
    ci = math.sqrt(2.0) / 2.0
    ct = O.transmission_cos(1.5, ci)
    assert abs(math.sqrt(1 - ci ** 2) / math.sqrt(1 - ct ** 2) - 1.5) < 1e-4      # 657-669
    assert abs(O.fresnel(1.33, 1.0)[2] - ((1 - 1.33) / 2.33) ** 2) < 1e-4          # 672-679
    assert abs(O.reflected_intensity(1.33, 0.0, 1.0, 550.0)) < 1e-6               # 682-686
    for t in (100.0, 200.0, 300.0, 400.0, 500.0):                                  # 689-700
        assert max(O.interference_color(1.33, t, 1.0)) > 0.01
    a, b = O.interference_color(1.33, 100.0, 1.0), O.interference_color(1.33, 300.0, 1.0)
    assert np.sum(np.abs(a - b)) > 0.01                                            # 703-716
    a, b = O.interference_color(1.33, 300.0, 1.0), O.interference_color(1.33, 300.0, 0.5)
    assert np.sum(np.abs(a - b)) > 0.001                                           # 719-731
    t1 = O.lut_1d(1.33, 0.0, 1000.0, 100, 1.0)
    assert t1.shape == (100, 3) and np.all((t1 >= 0) & (t1 <= 1))                 # 734-745
    assert O.lut_2d(1.33, 0.0, 500.0, 10, 5).shape == (10, 5, 3)                   # 748-755
    lin = O.interference_color(1.33, 300.0, 1.0, gamma=False)
    gam = O.interference_color(1.33, 300.0, 1.0)
    if 0.01 < lin[0] < 0.99:
        assert gam[0] >= lin[0]                                                    # 758-769
    cw = O.interference_color(1.33, 300.0, 1.0, lambdas=(700.0, 550.0, 400.0))
    assert np.all((cw >= 0) & (cw <= 1))                                           # 772-780


def test_fresnel_textbook(oracle):
    for n in (1.33, 1.5):
        for ci in np.linspace(0.05, 0.99, 40):
            ti = math.acos(ci)
            tt = math.asin(math.sin(ti) / n)
            rs_t = (math.sin(ti - tt) / math.sin(ti + tt)) ** 2
            rp_t = (math.tan(ti - tt) / math.tan(ti + tt)) ** 2
            rs, rp, ra = oracle.fresnel(n, ci)
            assert abs(rs - rs_t) <= 1e-12 and abs(rp - rp_t) <= 1e-12
            assert abs(ra - 0.5 * (rs_t + rp_t)) <= 1e-12
    assert oracle.fresnel(1.33, 0.0)[2] == 1.0


def test_airy_special_phases(oracle):
    n = 1.33
    R0 = ((1 - n) / (1 + n)) ** 2
    for lam in (450.0, 532.0, 650.0):
        for m in (1, 2, 5):
            d0 = (m - 0.5) * lam / (2 * n)
            assert oracle.reflected_intensity(n, d0, 1.0, lam) <= 1e-20
            d1 = m * lam / (2 * n)
            assert abs(oracle.reflected_intensity(n, d1, 1.0, lam) - 4 * R0 / (1 + R0) ** 2) <= 1e-14
    lut = oracle.lut_2d(n, 0.0, 1000.0, 64, 9)
    assert np.all(lut[0] == 0.0) and np.all(lut[1:, 0] == 1.0)


def _numpy_runtime_lut(n_film, intensity):
    """Independent float32 numpy restatement of interference_lut.rs:18-163."""
    f = np.float32
    t = (np.arange(256, dtype=f) / f(255)) * f(2000)
    c = np.arange(64, dtype=f) / f(63)
    T, Cc = np.meshgrid(t, c)
    sin_i = np.sqrt(np.maximum(f(1) - Cc * Cc, f(0)))
    sin_t = sin_i / f(n_film)
    cos_t = np.sqrt(np.maximum(f(1) - sin_t * sin_t, f(0)))
    rs = (Cc - f(n_film) * cos_t) / np.maximum(Cc + f(n_film) * cos_t, f(1e-4))
    rp = (f(n_film) * Cc - cos_t) / np.maximum(f(n_film) * Cc + cos_t, f(1e-4))
    R = (rs * rs + rp * rp) * f(0.5)
    op = f(2) * f(n_film) * T * cos_t
    g = lambda x, mu, s: np.exp(f(-0.5) * ((x - f(mu)) / f(s)) ** 2)
    xyz = np.zeros((3,) + T.shape, f)
    for w in map(f, (400, 450, 500, 550, 600, 650, 700)):
        ph = f(2) * f(np.pi) * op / w + f(np.pi)
        omr = np.maximum(f(1) - R, f(1e-3))
        F = f(4) * R / (omr * omr)
        s2 = np.sin(ph * f(0.5)) ** 2
        a = F * s2 / np.maximum(f(1) + F * s2, f(1e-3))
        cie = [max(f(1.056) * g(w, 599.8, 37.9) + f(0.362) * g(w, 442.0, 16.0) - f(0.065) * g(w, 501.1, 20.4), f(0)),
               max(f(0.821) * g(w, 568.8, 46.9) + f(0.286) * g(w, 530.9, 31.1), f(0)),
               max(f(1.217) * g(w, 437.0, 11.8) + f(0.681) * g(w, 459.0, 26.0), f(0))]
        for k in range(3):
            xyz[k] += f(cie[k]) * a
    xyz /= f(7)
    M = np.array([[3.2404542, -1.5371385, -0.4985314],
                  [-0.969266, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]], np.float32)
    rgb = np.einsum("ck,khw->hwc", M, xyz) * f(intensity)
    return rgb


def test_runtime_lut_numpy(oracle):
    for n, inten in ((1.33, 1.0), (1.45, 0.8)):
        want = _numpy_runtime_lut(n, inten)
        got = oracle.lut_rgb_f32(n, inten)
        assert np.max(np.abs(got - want)) <= 2e-5


def _fresnel_textbook(n, cos_i):
    """Unpolarised reflectance from the sin/tan Fresnel forms (FP64)."""
    ti = math.acos(cos_i)
    tt = math.asin(math.sin(ti) / n)
    if ti == 0.0:
        return ((1 - n) / (1 + n)) ** 2
    rs = (math.sin(ti - tt) / math.sin(ti + tt)) ** 2
    rp = (math.tan(ti - tt) / math.tan(ti + tt)) ** 2
    return 0.5 * (rs + rp)


LUT_W = (400.0, 450.0, 500.0, 550.0, 600.0, 650.0, 700.0)


@pytest.mark.parametrize("n", [1.33, 1.5])
def test_lut_zero_thickness_column(oracle, n):
    """At t = 0 every phase is pi, sin(pi/2) = 1, so each Airy term is
    F/(1+F) = 4R/(1+R)^2; the colour then depends on the angle only through
    that factor, so column ratios across angles equal ratios of 4R/(1+R)^2."""
    rows = np.arange(1, 64, dtype=np.float32) / np.float32(63)
    want = []
    for c in rows:
        _, r, f, a = oracle.lut_airy_terms(0.0, c, n)
        R = _fresnel_textbook(n, float(c))
        assert abs(float(r) - R) <= 2e-6 * R + 1e-7, (c, r, R)
        w = 4 * R / (1 + R) ** 2
        np.testing.assert_allclose(a, w, rtol=4e-6)
        want.append(w)
    want = np.array(want)
    lut = oracle.lut_rgb_f32(n, 1.0)
    col = lut[1:, 0, :].astype(np.float64)          # [angle][rgb]
    ratio = col / col[-1]
    np.testing.assert_allclose(ratio, np.repeat((want / want[-1])[:, None], 3, 1), rtol=2e-5)


def test_lut_grazing_row(oracle):
    """cos = 0: both Fresnel amplitudes are -1, R = 1, the clamp max(1-R,
    1e-3) gives F = 4e6, and a_k = F s^2/(1+F s^2), s = cos(pi OPD/lambda)."""
    n = 1.33
    f32 = np.float32
    for t in (np.arange(256, dtype=f32) / f32(255)) * f32(2000):
        cos_t, r, f, a = oracle.lut_airy_terms(t, 0.0, n)
        assert r == 1.0
        assert abs(float(f) - 4e6) <= 4e6 * 1e-6
        assert abs(float(cos_t) - math.sqrt(1 - 1 / n ** 2)) <= 1e-7
        op = 2 * n * float(t) * math.sqrt(1 - 1 / n ** 2)
        for k, w in enumerate(LUT_W):
            s = math.cos(math.pi * op / w)
            want = 4e6 * s * s / (1 + 4e6 * s * s)
            # FP32 phase: |d a / d phase| <= F |s c| / (1+F s^2)^2, phase error
            # ~ 1e-6 |phase| from FP32 rounding of the optical path and phase.
            phase = 2 * math.pi * op / w + math.pi
            sens = 4e6 * abs(s * math.sin(math.pi * op / w)) / (1 + 4e6 * s * s) ** 2
            assert abs(float(a[k]) - want) <= 2e-6 + sens * 1e-6 * phase, (t, w, a[k], want)


def test_lut_airy_matches_reflected_intensity(oracle):
    """The LUT's finesse form F sin^2(phi/2)/(1 + F sin^2(phi/2)) with phase
    2 pi OPD/lambda + pi equals the FP64 calculate_reflected_intensity
    2R(1 - cos d)/(1 + R^2 - 2R cos d), d = 2 pi (OPD + lambda/2)/lambda,
    whose special phases test_airy_special_phases pins: the two are
    algebraically identical, so they agree to FP32 rounding."""
    n = 1.33
    rng = np.random.default_rng(12)
    worst = 0.0
    for _ in range(600):
        t = np.float32(rng.uniform(0.5, 2000.0))
        c = np.float32(rng.uniform(0.15, 1.0))
        _, r, f, a = oracle.lut_airy_terms(t, c, n)
        for k, w in enumerate(LUT_W):
            want = oracle.reflected_intensity(n, float(t), float(c), w)
            phase = 4 * math.pi * n * float(t) / w + math.pi
            tol = 2e-6 + 1e-5 * want + float(f) * phase * 1e-6
            err = abs(float(a[k]) - want)
            worst = max(worst, err / tol)
            assert err <= tol, (t, c, w, a[k], want)
    print("LUT Airy vs reflected_intensity: worst err/tol %.3f" % worst)


def test_lut_colour_stage(oracle):
    """rgb = B a for one fixed 3 x 7 matrix B (colour curves / 7 times the
    XYZ -> sRGB matrix), and 7 M_xyz B — with M_xyz the published sRGB (D65)
    -> XYZ matrix of IEC 61966-2-1 — recovers non-negative curves shaped like
    the CIE 1931 2-degree observer at 400..700 nm (the reference's Gaussian
    fits: peaks at the table's wavelengths, correlation >= 0.97)."""
    rng = np.random.default_rng(0)
    A, Y = [], []
    for _ in range(300):
        t = np.float32(rng.uniform(0, 2000))
        c = np.float32(rng.uniform(0, 1))
        A.append(oracle.lut_airy_terms(t, c, 1.33)[3].astype(np.float64))
        Y.append(oracle.lut_thin_film(t, c, 1.33, 1.0).astype(np.float64))
    A, Y = np.array(A), np.array(Y)
    B = np.linalg.lstsq(A, Y, rcond=None)[0]
    assert np.max(np.abs(A @ B - Y)) <= 1e-6
    m_xyz = np.array([[0.4124564, 0.3575761, 0.1804375],
                      [0.2126729, 0.7151522, 0.0721750],
                      [0.0193339, 0.1191920, 0.9503041]])
    curves = 7 * (m_xyz @ B.T)
    cie1931 = np.array([[0.01431, 0.33620, 0.00490, 0.43345, 1.06220, 0.28350, 0.01136],
                        [0.000396, 0.03800, 0.32300, 0.99495, 0.63100, 0.10700, 0.004102],
                        [0.06785, 1.77211, 0.27200, 0.00875, 0.00080, 0.00002, 0.0]])
    assert np.all(curves >= -1e-5)
    for k in range(3):
        assert np.argmax(curves[k]) == np.argmax(cie1931[k])
        assert np.corrcoef(curves[k], cie1931[k])[0, 1] >= 0.97
    # Intensity scales the colour linearly.
    np.testing.assert_allclose(oracle.lut_thin_film(321.0, 0.7, 1.33, 0.5),
                               0.5 * oracle.lut_thin_film(321.0, 0.7, 1.33, 1.0), rtol=1e-6, atol=1e-9)


def test_spectral_one_line_curves(oracle):
    """Curves that are 1 at one wavelength and 0 elsewhere turn the
    81-wavelength contraction into I(lambda_k)/81, I the phase-pinned FP64
    reflected intensity at the cell's view angle (away from grazing views,
    where the finesse clamp max(1 - R, 1e-3) of the LUT form applies)."""
    nt, npf, R = 21, 40, 0.025
    rng = np.random.default_rng(8)
    h = rng.uniform(50e-9, 3000e-9, (nt, npf))
    cam = (0.03, 0.06, 0.08)
    n = 1.33
    for ks in ((0, 40, 80), (13, 27, 66)):
        cmf = np.zeros((81, 3))
        for c, k in enumerate(ks):
            cmf[k, c] = 1.0
        rgb = oracle.spectral_rgb(h, R, cam, n, cmf)
        checked = 0
        for i in range(nt):
            for j in range(npf):
                cv = oracle.view_cos(i * math.pi / (nt - 1), j * 2 * math.pi / npf, R, cam)
                if 1.0 - oracle.fresnel(n, cv)[2] <= 2e-3:
                    continue
                for c, k in enumerate(ks):
                    lam = 380.0 + 5.0 * k
                    want = oracle.reflected_intensity(n, h[i, j] * 1e9, cv, lam) / 81.0
                    ra = oracle.fresnel(n, cv)[2]
                    F = 4 * ra / (1 - ra) ** 2
                    phase = 4 * math.pi * n * h[i, j] * 1e9 / lam + math.pi
                    assert abs(rgb[c, i, j] - want) <= 1e-14 + F * phase * 1e-15, (i, j, k)
                checked += 1
        assert checked > nt * npf // 2


def test_reference_lut_asserts(oracle):
    lut = oracle.lut_rgba8(1.33, 1.0)
    assert lut.size == 256 * 64 * 4 and np.all(lut[..., 3] == 255)          # 170-174
    rgb = oracle.lut_thin_film(500.0, 0.9, 1.33, 1.0)                         # 177-185
    assert np.all((rgb >= 0) & (rgb <= 1)) and rgb.sum() > 0.01
    # snells_law / fresnel_unpolarized asserts (188-204) hold in the LUT's
    # cos = 1 row: R0 = ((1-1.33)/2.33)^2 in (0.01, 0.05).
    R0 = ((1 - 1.33) / 2.33) ** 2
    assert 0.01 < R0 < 0.05
    pre = np.clip(oracle.lut_rgb_f32(1.33, 1.0), 0, 1)
    np.testing.assert_array_equal(lut[..., :3], np.floor(pre * np.float32(255)).astype(np.uint8))


def test_spectral_reduces_to_lut(oracle):
    """A colour curve that is non-zero only at 400, 450, ..., 700 nm and holds
    81/7 * (XYZ->RGB matrix . CIE fits) there turns the 81-wavelength sum into
    the runtime LUT's 7-sample colour at the cell's view angle."""
    M = np.array([[3.2404542, -1.5371385, -0.4985314],
                  [-0.969266, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]])
    g = lambda x, mu, s: math.exp(-0.5 * ((x - mu) / s) ** 2)
    cmf = np.zeros((81, 3))
    for w in range(400, 701, 50):
        cie = np.array([max(1.056 * g(w, 599.8, 37.9) + 0.362 * g(w, 442, 16) - 0.065 * g(w, 501.1, 20.4), 0),
                        max(0.821 * g(w, 568.8, 46.9) + 0.286 * g(w, 530.9, 31.1), 0),
                        max(1.217 * g(w, 437, 11.8) + 0.681 * g(w, 459, 26), 0)])
        cmf[(w - 380) // 5] = 81.0 / 7.0 * (M @ cie)
    nt, npf, R = 17, 32, 0.025
    rng = np.random.default_rng(5)
    h = rng.uniform(100e-9, 900e-9, (nt, npf))
    cam = (0.03, 0.06, 0.08)
    rgb = oracle.spectral_rgb(h, R, cam, 1.33, cmf)
    for i in range(1, nt - 1):
        for j in range(0, npf, 5):
            cv = oracle.view_cos(i * math.pi / (nt - 1), j * 2 * math.pi / npf, R, cam)
            want = oracle.lut_thin_film(h[i, j] * 1e9, cv, 1.33, 1.0)
            assert np.max(np.abs(rgb[:, i, j] - want)) <= 2e-5


def test_view_cos(oracle):
    R = 0.025
    cam = (0.0, 0.0, 4 * R)
    # the point facing the camera sees it head-on
    assert abs(oracle.view_cos(math.pi / 2, math.pi / 2, R, cam) - 1.0) <= 1e-15
    # on the silhouette circle (n . (cam - p) = 0) the view is grazing
    ang = math.acos(R / (4 * R))
    th = math.pi / 2
    ph = math.pi / 2 - ang
    assert oracle.view_cos(th, ph, R, cam) <= 1e-12
