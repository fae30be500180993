"""GPU parity of the interference kernels.

Bars (tolerances stated per test):
  * LUT addressing (thickness / cos of every texel) is formed on the host in
    FP64 like the reference, so texel coordinates match bit for bit;
  * RGB within 1e-3 of the FP64 oracle;
  * the RGBA8 LUT is compared before quantisation (one LSB = 3.9e-3), and the
    quantised bytes may differ by one LSB only where the value sits within
    1e-5 of a quantisation boundary.
"""
import ctypes as C
import math

import numpy as np
import pytest

from soap_bubble_simulation_amd import _lib as L
from soap_bubble_simulation_amd import workloads as W
from soap_bubble_simulation_amd.drainage import DrainageContext
from soap_bubble_simulation_amd.interference import (InterferenceCalculator, RgbColor,
                                                     generate_interference_lut,
                                                     interference_lut_rgb, interference_lut_rgba8,
                                                     spectral_rgb)

pytestmark = pytest.mark.gpu


def test_runtime_lut_prequant(oracle):
    got = interference_lut_rgb(1.33, 1.0)
    want = oracle.lut_rgb_f32(1.33, 1.0)
    err = np.max(np.abs(got.astype(np.float64) - want))
    print("runtime LUT max |rgb - oracle| =", err)
    assert err <= 2e-5


def test_runtime_lut_rgba8(oracle):
    lut = generate_interference_lut(1.33, 1.0)
    assert len(lut) == 256 * 64 * 4                      # interference_lut.rs:170-174
    got = np.frombuffer(lut, np.uint8).reshape(64, 256, 4)
    want = oracle.lut_rgba8(1.33, 1.0)
    assert np.all(got[..., 3] == 255)
    diff = np.abs(got.astype(int) - want.astype(int))
    assert diff.max() <= 1
    pre = np.clip(oracle.lut_rgb_f32(1.33, 1.0), 0, 1) * 255.0
    near = np.abs(pre - np.round(pre)) <= 255.0 * 1e-5
    assert np.all(near[diff[..., :3] > 0])
    print("RGBA8 bytes differing by one LSB:", int((diff > 0).sum()))


def test_runtime_lut_intensity_and_index(oracle):
    for n, inten in ((1.40, 0.7), (1.25, 1.3)):
        got = interference_lut_rgb(n, inten)
        want = oracle.lut_rgb_f32(n, inten)
        assert np.max(np.abs(got - want)) <= 2e-5 * max(1.0, inten)


def test_lut_reference_unit_tests():
    """interference_lut.rs:177-204 through the device LUT kernel."""
    lut = interference_lut_rgb(1.33, 1.0)
    # thin_film_interference(500 nm, 0.9): texel t = 500/2000*255 is not on
    # the grid, so check the neighbouring texels stay in [0,1] with colour.
    a = int(round(0.9 * 63))
    for t in (63, 64):
        rgb = lut[a, t]
        assert np.all(rgb >= -0.05) and np.all(rgb <= 1.05)
    assert lut[a, 64].sum() > 0.01


def test_lut_2d_default_grid(oracle):
    """1024 thicknesses (0-1000 nm) x 90 angles at 650/532/450 nm."""
    calc = InterferenceCalculator(1.33)
    got = calc.generate_2d_color_lookup_table_array(0.0, 1000.0, 1024, 90)
    want = oracle.lut_2d(1.33, 0.0, 1000.0, 1024, 90)
    err = np.max(np.abs(got - want))
    print("2-D LUT max |rgb - oracle| =", err)
    assert err <= 1e-3
    # Closed forms: d = 0 row is black; the cos = 0 column has R = 1, I = 1.
    assert np.all(got[0] == 0.0)
    assert np.all(got[1:, 0] == 1.0)
    lin = calc.generate_2d_color_lookup_table_array(0.0, 1000.0, 64, 16, gamma=False)
    assert np.max(np.abs(lin - oracle.lut_2d(1.33, 0.0, 1000.0, 64, 16, gamma=False))) <= 1e-4


def test_lut_1d(oracle):
    calc = InterferenceCalculator(1.33)
    got = calc.generate_color_lookup_table_array(0.0, 1000.0, 100, 1.0)
    want = oracle.lut_1d(1.33, 0.0, 1000.0, 100, 1.0)
    assert np.max(np.abs(got - want)) <= 1e-3
    table = calc.generate_color_lookup_table(0.0, 1000.0, 100, 1.0)   # interference.rs:734-745
    assert len(table) == 100
    assert all(0 <= c.red <= 1 and 0 <= c.green <= 1 and 0 <= c.blue <= 1 for c in table)


def test_interference_closed_forms():
    calc = InterferenceCalculator(1.33)
    n = 1.33
    R0 = ((1 - n) / (1 + n)) ** 2
    for lam in (450.0, 532.0, 650.0):
        for m in (1, 2, 3):
            # delta = 2 pi m: OPD = 2 n d + lam/2 = m lam  ->  I = 0
            d0 = (m - 0.5) * lam / (2 * n)
            assert calc.calculate_reflected_intensity(d0, 1.0, lam) <= 1e-9
            # delta = (2m+1) pi: OPD = (m + 1/2) lam  ->  I = 4R/(1+R)^2
            d1 = m * lam / (2 * n)
            v = calc.calculate_reflected_intensity(d1, 1.0, lam)
            assert abs(v - 4 * R0 / (1 + R0) ** 2) <= 1e-6
    f = calc.calculate_fresnel_reflection(1.0)
    assert abs(f.reflectance_average - R0) <= 1e-6
    assert calc.calculate_fresnel_reflection(0.0).reflectance_average == 1.0


def test_reference_interference_unit_tests():
    """interference.rs:619-780 through the GPU-backed calculator."""
    c = RgbColor.new(1.5, -0.5, 0.5)
    assert (c.red, c.green, c.blue) == (1.0, 0.0, 0.5)
    assert RgbColor.new(1.0, 0.5, 0.0).to_u8_tuple() == (255, 128, 0)
    calc = InterferenceCalculator.new(1.33)
    assert abs(calc.refractive_index_film() - 1.33) < 1e-6
    assert abs(calc.refractive_index_medium() - 1.0) < 1e-6
    assert abs(calc.calculate_transmission_angle_cos(1.0) - 1.0) < 1e-6
    c15 = InterferenceCalculator.new(1.5)
    ci = math.sqrt(2.0) / 2.0
    ct = c15.calculate_transmission_angle_cos(ci)
    ratio = math.sqrt(1 - ci ** 2) / math.sqrt(1 - ct ** 2)
    assert abs(ratio - 1.5) < 1e-4
    expected = ((1.0 - 1.33) / (1.0 + 1.33)) ** 2
    assert abs(calc.calculate_fresnel_reflection(1.0).reflectance_average - expected) < 1e-4
    d = InterferenceCalculator.default_soap_film()
    assert abs(d.calculate_reflected_intensity(0.0, 1.0, 550.0)) < 1e-6
    for t in (100.0, 200.0, 300.0, 400.0, 500.0):
        col = d.calculate_interference_color(t, 1.0)
        assert max(col.red, col.green, col.blue) > 0.01
    a, b = d.calculate_interference_color(100.0, 1.0), d.calculate_interference_color(300.0, 1.0)
    assert abs(a.red - b.red) + abs(a.green - b.green) + abs(a.blue - b.blue) > 0.01
    a, b = d.calculate_interference_color(300.0, 1.0), d.calculate_interference_color(300.0, 0.5)
    assert abs(a.red - b.red) + abs(a.green - b.green) + abs(a.blue - b.blue) > 0.001
    t2 = d.generate_2d_color_lookup_table(0.0, 500.0, 10, 5)
    assert len(t2) == 10 and len(t2[0]) == 5
    lin = d.calculate_interference_color_linear(300.0, 1.0)
    gam = d.calculate_interference_color(300.0, 1.0)
    if 0.01 < lin.red < 0.99:
        assert gam.red >= lin.red
    cw = InterferenceCalculator.with_custom_wavelengths(1.33, 700.0, 550.0, 400.0)
    col = cw.calculate_interference_color(300.0, 1.0)
    assert 0 <= col.red <= 1 and 0 <= col.green <= 1 and 0 <= col.blue <= 1


def test_point_evaluations_match_oracle(oracle):
    rng = np.random.default_rng(4)
    n = 4000
    d = rng.uniform(0, 2000, n).astype(np.float32)
    c = rng.uniform(0, 1, n).astype(np.float32)
    lam = rng.uniform(380, 780, n).astype(np.float32)
    calc = InterferenceCalculator(1.33)
    got = calc.reflected_intensity(d, c, lam)
    want = np.array([oracle.reflected_intensity(1.33, float(a), float(b), float(w))
                     for a, b, w in zip(d, c, lam)])
    assert np.max(np.abs(got - want)) <= 1e-4
    terms = calc.terms(d[:200], c[:200], lam[:200])
    for k in range(200):
        ct = oracle.transmission_cos(1.33, float(c[k]))
        assert abs(terms[k, 0] - ct) <= 1e-6
        rs, rp, ra = oracle.fresnel(1.33, float(c[k]))
        assert abs(terms[k, 5] - ra) <= 1e-6
    cols = calc.colors(d[:500], c[:500], gamma=True)
    want_c = np.array([oracle.interference_color(1.33, float(a), float(b)) for a, b in zip(d[:500], c[:500])])
    assert np.max(np.abs(cols - want_c)) <= 1e-3
    opd = calc.calculate_optical_path_difference(300.0, 0.8, 532.0)
    assert abs(opd - oracle.optical_path_difference(1.33, 300.0, 0.8, 532.0)) <= 1e-3
    ph = calc.calculate_phase_difference(opd, 532.0)
    assert abs(ph - oracle.phase_difference(opd, 532.0)) <= 1e-5


def _spectral_case(nt, npf, seed):
    h0 = W.smooth_bubble(nt, npf, seed)
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    ctx = DrainageContext(p, nt, npf, 500e-9)
    ctx.set_fields(h0, np.full_like(h0, 0.5))
    cam = W.camera(p.bubble_radius)
    cmf = W.spectral_cmf(seed)
    return ctx, h0, p, cam, cmf


def test_spectral_small_grid(oracle):
    nt, npf = 96, 192
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 12)
    got = spectral_rgb(ctx, cam, cmf.astype(np.float32))
    want = oracle.spectral_rgb(h0, p.bubble_radius, cam, 1.33, cmf.astype(np.float32).astype(np.float64))
    err = np.max(np.abs(got - want))
    print("spectral 96x192 max |rgb - oracle| =", err)
    assert err <= 1e-3
    assert np.max(np.abs(want)) > 0.05


@pytest.mark.parametrize("nt,npf", [(7, 13), (5, 30), (9, 66)])
def test_spectral_ragged_grids(oracle, nt, npf):
    """Cell counts that are not multiples of 4 or 64 (91, 150, 594 cells):
    partial 64-cell groups and the scalar RGB stores."""
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 31)
    cm = cmf.astype(np.float32)
    got = spectral_rgb(ctx, cam, cm)
    want = oracle.spectral_rgb(h0, p.bubble_radius, cam, 1.33, cm.astype(np.float64))
    assert got.shape == want.shape == (3, nt, npf)
    err = np.max(np.abs(got - want))
    print("spectral %dx%d max |rgb - oracle| = %.3g" % (nt, npf, err))
    assert err <= 1e-3


def test_spectral_curves_not_exact_in_bf16(oracle):
    """Curves with a nonzero BF16 remainder take the third (hi x curve-lo) MFMA."""
    nt, npf = 64, 128
    ctx, h0, p, cam, _ = _spectral_case(nt, npf, 5)
    cmf = W.spectral_cmf(5, bf16=False).astype(np.float32)
    assert np.any(cmf.view(np.uint32) & 0xFFFF)
    got = spectral_rgb(ctx, cam, cmf)
    want = oracle.spectral_rgb(h0, p.bubble_radius, cam, 1.33, cmf.astype(np.float64))
    err = np.max(np.abs(got - want))
    print("spectral, FP32 curves, max |rgb - oracle| =", err)
    assert err <= 1e-3


@pytest.mark.parametrize("scale", [1.0, 20.0])
def test_spectral_operand_formats(oracle, scale):
    """Both operand formats against FP64 on the same curves.  The kernel
    contracts r / 81, which FP16 rounds to nearest within 2^-18 on
    [0, 1/81], so its RGB error is bounded by 2^-18 sum_k |cmf[k][c]| (plus
    FP32 rounding); the automatic choice takes FP16 while that bound is at
    most 2.5e-4 and the BF16 hi/lo pair past it (here: the curves scaled by
    20)."""
    nt, npf = 96, 192
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 14)
    cm = (cmf * scale).astype(np.float32)
    want = oracle.spectral_rgb(h0, p.bubble_radius, cam, 1.33, cm.astype(np.float64)).reshape(3, -1)
    bound16 = 2.0 ** -18 * np.abs(cm.astype(np.float64)).sum(axis=0)
    auto = L.SPECTRAL_OPERAND_FP16 if bound16.max() <= 2.5e-4 else L.SPECTRAL_OPERAND_BF16_PAIR
    assert auto == (L.SPECTRAL_OPERAND_FP16 if scale == 1.0 else L.SPECTRAL_OPERAND_BF16_PAIR)
    lib = L.lib()
    got = {}
    for kind in (L.SPECTRAL_OPERAND_BF16_PAIR, L.SPECTRAL_OPERAND_FP16, L.SPECTRAL_OPERAND_AUTO):
        L.check(lib.sf_spectral_set_operand(ctx.handle, kind))
        got[kind] = spectral_rgb(ctx, cam, cm).reshape(3, -1)
        used = C.c_int32(-1)
        L.check(lib.sf_spectral_get_operand(ctx.handle, C.byref(used)))
        assert used.value == (auto if kind == L.SPECTRAL_OPERAND_AUTO else kind)
    assert np.array_equal(got[L.SPECTRAL_OPERAND_AUTO], got[auto])
    err16 = np.max(np.abs(got[L.SPECTRAL_OPERAND_FP16] - want), axis=1)
    errbf = np.max(np.abs(got[L.SPECTRAL_OPERAND_BF16_PAIR] - want), axis=1)
    print("scale %g: FP16 max err per channel %s (bound %s), BF16 pair %s" % (scale, err16, bound16, errbf))
    assert np.all(err16 <= bound16 + 2e-5 * scale)
    assert np.all(errbf <= 1e-4 * scale)
    assert np.max(errbf) <= 1e-3 and np.max(np.abs(got[auto] - want)) <= 1e-3
    L.check(lib.sf_spectral_set_operand(ctx.handle, L.SPECTRAL_OPERAND_AUTO))


def test_spectral_trailing_slots(oracle):
    """Trailing wavelength slots (slot 20: 780 nm and padding; slot 19:
    760-775 nm) are evaluated only where the curves weigh them: they are
    skipped while the skipped wavelengths' sum_k |cmf[k][c]| / 81 stays
    within 1e-5 for every c, which bounds the RGB change.  Curves that weigh
    them take the slots; skipping them there would move the RGB by 1e-2."""
    nt, npf = 64, 128
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 9)
    base = cmf.astype(np.float32)
    cases = []
    heavy = base.copy()
    heavy[76:81] = [0.5, 0.3, 0.2]                 # both slots weigh: none skipped
    cases.append(("both slots weighed", heavy, 0.0))
    cie = base.copy()
    cie[76:80] = [0.3, 0.1, 0.05]
    cie[80] = [4.2e-5, 1.5e-5, 0.0]               # CIE 1931 at 780 nm: slot 20 skipped
    cases.append(("780 nm light", cie, 1e-5))
    short = base.copy()
    short[76:81] = [1e-5, 2e-6, 0.0]               # curves ending below 760 nm: slots 19-20 skipped
    cases.append(("760-780 nm light", short, 1e-5))
    for name, cm, tail in cases:
        got = spectral_rgb(ctx, cam, cm).reshape(3, -1)
        want = oracle.spectral_rgb(h0, p.bubble_radius, cam, 1.33, cm.astype(np.float64)).reshape(3, -1)
        bound16 = 2.0 ** -18 * np.abs(cm.astype(np.float64)).sum(axis=0)
        err = np.max(np.abs(got - want), axis=1)
        print("%s: max err per channel %s (bound %s)" % (name, err, bound16 + tail + 2e-5))
        assert np.all(err <= bound16 + tail + 2e-5), name


def test_spectral_fp16_range():
    """Curves past FP16's range: AUTO takes the BF16 pair, a forced FP16
    format is an error at the spectral call."""
    nt, npf = 32, 64
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 2)
    cm = cmf.astype(np.float32).copy()
    cm[40, 1] = 3.0e4
    lib = L.lib()
    rgb = spectral_rgb(ctx, cam, cm)
    assert np.all(np.isfinite(rgb))
    used = C.c_int32(-1)
    L.check(lib.sf_spectral_get_operand(ctx.handle, C.byref(used)))
    assert used.value == L.SPECTRAL_OPERAND_BF16_PAIR
    L.check(lib.sf_spectral_set_operand(ctx.handle, L.SPECTRAL_OPERAND_FP16))
    with pytest.raises(L.SoapFilmError, match="FP16"):
        spectral_rgb(ctx, cam, cm)
    # The stepping call refuses before it queues any step.
    t0 = ctx.stats().current_time
    camera = L.Camera()
    for k in range(3):
        camera.position[k] = cam[k]
    assert lib.sf_drainage_step_spectral(ctx.handle, 3, 0.0, C.byref(camera), 1.33, L.fptr(cm)) != 0
    assert "FP16" in L.last_error()
    assert ctx.stats().current_time == t0
    assert lib.sf_spectral_set_operand(ctx.handle, 7) != 0


def test_spectral_thick_film(oracle):
    """Optical paths past 5e4 nm (films of 20-60 um) take the range-reduced
    square form in every wave."""
    nt, npf = 48, 96
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 6)
    h = h0 * 60.0
    ctx.set_fields(h, np.full_like(h, 0.5))
    cm = cmf.astype(np.float32)
    got = spectral_rgb(ctx, cam, cm)
    want = oracle.spectral_rgb(h, p.bubble_radius, cam, 1.33, cm.astype(np.float64))
    err = np.max(np.abs(got - want))
    print("spectral, 20-60 um films, max |rgb - oracle| =", err)
    assert err <= 1e-3
    assert np.max(np.abs(want)) > 0.05


def test_spectral_full_grid_sampled(oracle):
    """2048 x 4096 (the benchmark grid): 20000 sampled cells against FP64."""
    nt, npf = 2048, 4096
    ctx, h0, p, cam, cmf = _spectral_case(nt, npf, 3)
    cm = cmf.astype(np.float32)
    got = spectral_rgb(ctx, cam, cm)
    rng = np.random.default_rng(8)
    cells = np.concatenate([rng.choice(nt * npf, 20000, replace=False),
                            np.arange(npf), (nt - 1) * npf + np.arange(npf)])
    want = oracle.spectral_cells(h0, cells, p.bubble_radius, cam, 1.33, cm.astype(np.float64))
    g = got.reshape(3, -1)[:, cells].T
    err = np.max(np.abs(g - want))
    print("spectral 2048x4096 sampled max |rgb - oracle| =", err)
    assert err <= 1e-3


@pytest.mark.parametrize("nt,npf,steps,flags", [
    (96, 192, 7, L.FLAG_TFE | L.FLAG_SURFACTANT | L.FLAG_DIAG),
    (512, 1024, 20, L.FLAG_TFE | L.FLAG_SURFACTANT | L.FLAG_DIAG | L.FLAG_MARANGONI),
])
def test_step_spectral_matches_serial(oracle, nt, npf, steps, flags):
    """sf_drainage_step_spectral gives bit-identical state and RGB to the
    sequence of separate calls {sf_drainage_step, sf_spectral_rgb}; rebases
    fall inside the run.  The RGB is then checked against FP64 on the final
    state."""
    h0 = W.smooth_bubble(nt, npf, 21)
    cmf = W.spectral_cmf(21).astype(np.float32)
    cam_xyz = None
    ctxs = []
    for _ in range(2):
        p = L.Params()
        L.check(L.lib().sf_params_cpu_default(p))
        p.flags = flags
        ctx = DrainageContext(p, nt, npf, 500e-9)
        ctx.set_fields(h0, np.full_like(h0, 0.5))
        ctxs.append(ctx)
        cam_xyz = W.camera(p.bubble_radius)
    a, b = ctxs
    dt = a.stats().dt_cfl * a.params().cfl_safety
    cam = L.Camera()
    for k in range(3):
        cam.position[k] = cam_xyz[k]
    lib = L.lib()
    L.check(lib.sf_drainage_step_spectral(a.handle, steps, dt, C.byref(cam), 1.33, L.fptr(cmf)))
    rgb_a = np.empty((3, nt, npf), np.float32)
    L.check(lib.sf_spectral_copy_rgb(a.handle, L.fptr(rgb_a)))
    for _ in range(steps):
        b.step(1, dt)
        rgb_b = spectral_rgb(b, cam_xyz, cmf)
    ha, ga = a.fields()
    hb, gb = b.fields()
    assert np.array_equal(ha, hb) and np.array_equal(ga, gb)
    assert np.array_equal(rgb_a, rgb_b)
    rng = np.random.default_rng(4)
    cells = rng.choice(nt * npf, 3000, replace=False)
    want = oracle.spectral_cells(ha, cells, a.params().bubble_radius, cam_xyz, 1.33, cmf.astype(np.float64))
    err = np.max(np.abs(rgb_a.reshape(3, -1)[:, cells].T - want))
    print("step_spectral %dx%d after %d steps: max |rgb - oracle| = %.3g" % (nt, npf, steps, err))
    assert err <= 1e-3
