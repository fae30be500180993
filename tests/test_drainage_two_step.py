"""Two drainage steps per stencil launch (temporal blocking).

``drainage_band2_kernel`` streams a band once per two steps.  Its cell
update is the one-step kernels' (drainage.rs:298-392 interior step with the
surfactant transport of drainage.wgsl:178-201, the marginal-regeneration
fronts of drainage.rs:395-461), and the pole means between its two steps are
summed in the one-step order, so the bar is equality bit for bit with one
step per launch: every (hi, lo) pair, Γ, the pole rows, the front masks and
the simulation time.  The fused FP64 diagnostics of both steps are compared
too: the CFL step exactly (a maximum), the film volume within 1e-13 relative
(the same cell values summed over other block partitions).  Both contexts
take the float4 kernels (SOAPFILM_STENCIL=vec4) so that their column tiles,
and with them the pole partial sums, are the same.  Trajectories are also
compared with the FP64 oracle directly, at the bars of the one-step parity
tests, and the spectral frames of both states of a launch with the frames of
one-step launches.
"""
import numpy as np
import pytest

from soap_bubble_simulation_amd import _lib as L
from soap_bubble_simulation_amd import workloads as W
from soap_bubble_simulation_amd.drainage import DrainageContext

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def params(flags):
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    p.flags = flags
    return p


def run_pair(monkeypatch, nt, npf, flags, steps, rows2=None, dt=0.0, seed=7):
    """One-step and two-step contexts from the same smooth film, advanced
    by each count in ``steps`` (dt 0: the CFL step of the state at the start
    of the call, equal for both while their states are), equal after each."""
    monkeypatch.setenv("SOAPFILM_STENCIL", "vec4")
    if rows2 is not None:
        monkeypatch.setenv("SOAPFILM_ROWS_PER_BLOCK2", str(rows2))
    p = params(flags)
    h0 = W.smooth_bubble(nt, npf, seed=seed)
    rng = np.random.default_rng(seed)
    g0 = 0.5 + 0.05 * rng.standard_normal((nt, npf))
    one = DrainageContext(p, nt, npf, 500e-9)
    two = DrainageContext(p, nt, npf, 500e-9)
    one.set_steps_per_launch(1)
    two.set_steps_per_launch(2)
    assert one.steps_per_launch() == 1
    assert two.steps_per_launch() == 2, (nt, npf, flags)
    for c in (one, two):
        c.set_fields(h0, g0)
    hs, gs = one.fields()
    for n in steps:
        one.step(n, dt)
        two.step(n, dt)
        h1, g1 = one.fields()
        h2, g2 = two.fields()
        np.testing.assert_array_equal(h2, h1)
        np.testing.assert_array_equal(g2, g1)
        assert two.current_time() == one.current_time()
        if flags & L.FLAG_TFE:
            assert two.active_tfe_count() == one.active_tfe_count()
            np.testing.assert_array_equal(two.tfe_mask(), one.tfe_mask())
    if flags & L.FLAG_DIAG:
        m1, c1 = one.history()
        m2, c2 = two.history()
        assert len(m1) == len(m2) == sum(steps)
        np.testing.assert_array_equal(c2, c1)
        assert np.max(np.abs(m2 - m1) / np.abs(m1)) <= 1e-13
    # the comparison has teeth: the film and, when transported, Γ moved
    assert rel_l2(h1 - hs, hs) > 1e-9
    if flags & L.FLAG_SURFACTANT:
        assert np.max(np.abs(g1 - gs)) > 1e-6
    return one, two, h0


HEADLINE_FLAGS = L.FLAG_TFE | L.FLAG_SURFACTANT | L.FLAG_DIAG


@pytest.mark.parametrize("flags", [L.FLAG_SURFACTANT, 0, L.FLAG_SURFACTANT | L.FLAG_POLES_COPY_H,
                                   L.FLAG_POLES_COPY_H, HEADLINE_FLAGS, L.FLAG_TFE | L.FLAG_DIAG,
                                   L.FLAG_DIAG | L.FLAG_POLES_COPY_H])
@pytest.mark.parametrize("nt,npf", [(130, 512), (66, 1280), (34, 768), (4, 256), (257, 2048)])
def test_two_step_bit_identical(monkeypatch, flags, nt, npf):
    """Odd and even step counts, across the rebase period (8), on grids with
    one to several bands, partial column tiles (1280 = 5 waves of 256 in
    two four-wave tiles), three-wave blocks (768) and a two-row interior."""
    run_pair(monkeypatch, nt, npf, flags, steps=(1, 2, 5, 8, 13, 3))


@pytest.mark.parametrize("rows2", [1, 3, 16, 64])
def test_two_step_band_heights(monkeypatch, rows2):
    """Band heights from one row to the largest."""
    run_pair(monkeypatch, 130, 1024, L.FLAG_SURFACTANT, steps=(7, 10), rows2=rows2)
    run_pair(monkeypatch, 67, 512, 0, steps=(3, 6), rows2=rows2)


def test_two_step_headline_grid(monkeypatch):
    """2048 x 4096, the headline field, surfactant transport."""
    run_pair(monkeypatch, 2048, 4096, L.FLAG_SURFACTANT, steps=(4, 9), seed=2)


@pytest.mark.parametrize("flags", [HEADLINE_FLAGS, L.FLAG_TFE | L.FLAG_DIAG])
@pytest.mark.parametrize("nt,npf", [(66, 512), (130, 768), (34, 1024)])
def test_two_step_with_fronts(monkeypatch, flags, nt, npf):
    """CFL steps of about 0.03-0.1 s on these grids: fronts spawn after
    0.5 s, so the later launches thin the film inside them, in both steps
    and in the halo rows and edge columns the second step reads."""
    one, two, _ = run_pair(monkeypatch, nt, npf, flags, steps=(9, 12, 1, 17, 8, 25))
    assert one.active_tfe_count() > 0
    assert np.any(one.tfe_mask())


def test_two_step_ineligible_flags_keep_one_step(monkeypatch):
    monkeypatch.setenv("SOAPFILM_STENCIL", "vec4")
    for flags in (L.FLAG_MARANGONI, L.FLAG_SURFACTANT | L.FLAG_POLES_AVG_GAMMA):
        c = DrainageContext(params(flags), 66, 512, 500e-9)
        c.set_steps_per_launch(2)
        assert c.steps_per_launch() == 1, flags
    for flags in (L.FLAG_TFE, L.FLAG_DIAG, HEADLINE_FLAGS):
        c = DrainageContext(params(flags), 66, 512, 500e-9)
        c.set_steps_per_launch(2)
        assert c.steps_per_launch() == 2, flags
    c = DrainageContext(params(L.FLAG_SURFACTANT), 66, 500, 500e-9)     # n_phi % 256 != 0
    c.set_steps_per_launch(2)
    assert c.steps_per_launch() == 1
    with pytest.raises(L.SoapFilmError):
        c.set_steps_per_launch(3)


def test_two_step_oracle_trajectory(oracle, monkeypatch):
    """256 x 1024, surfactant transport, 200 steps two per launch against
    the FP64 oracle: h within 1e-6 and the increment within 1e-5 relative
    L2, Γ within 1e-6, rupture mask and time exact."""
    monkeypatch.setenv("SOAPFILM_STENCIL", "vec4")
    nt, npf = 256, 1024
    p = params(L.FLAG_SURFACTANT)
    h0 = W.smooth_bubble(nt, npf, seed=4)
    g0 = np.full_like(h0, 0.5)
    ctx = DrainageContext(p, nt, npf, 500e-9)
    ctx.set_steps_per_launch(2)
    assert ctx.steps_per_launch() == 2
    q = oracle.Params()
    for name, _ in oracle.Params._fields_:
        setattr(q, name, getattr(p, name))
    orc = oracle.Sim(q, nt, npf)
    ctx.set_fields(h0, g0)
    orc.set_fields(h0, g0)
    dt = orc.cfl_dt()
    ctx.step(200, dt)
    orc.step(dt, 200)
    hg, gg = ctx.fields()
    ho, go = orc.fields()
    assert rel_l2(hg, ho) <= 1e-6
    assert rel_l2(hg - h0, ho - h0) <= 1e-5, rel_l2(hg - h0, ho - h0)
    assert np.max(np.abs(gg - go)) <= 1e-6
    np.testing.assert_array_equal(ctx.rupture_mask(), orc.rupture_mask())
    assert ctx.current_time() == orc.current_time()


def test_two_step_headline_flags_oracle_trajectory(oracle, monkeypatch):
    """128 x 512, the headline flags (fronts, surfactant transport, fused
    diagnostics), 60 CFL steps (0.029 s each) two per launch against the
    FP64 oracle, fronts spawning at 0.5 s: h within 1e-6 and the increment
    within 1e-5 relative L2, Γ within 1e-6, rupture and front masks and time
    exact, the per-step FP64 film volume and CFL step within 1e-9 of the
    oracle's (the bar of the one-step history test)."""
    monkeypatch.setenv("SOAPFILM_STENCIL", "vec4")
    nt, npf = 128, 512
    p = params(HEADLINE_FLAGS)
    h0 = W.smooth_bubble(nt, npf, seed=5)
    g0 = np.full_like(h0, 0.5)
    ctx = DrainageContext(p, nt, npf, 500e-9)
    ctx.set_steps_per_launch(2)
    assert ctx.steps_per_launch() == 2
    q = oracle.Params()
    for name, _ in oracle.Params._fields_:
        setattr(q, name, getattr(p, name))
    orc = oracle.Sim(q, nt, npf)
    ctx.set_fields(h0, g0)
    orc.set_fields(h0, g0)
    dt = orc.cfl_dt()
    masses, cfls = [], []
    for _ in range(3):
        ctx.step(20, dt)
        for _ in range(20):
            orc.step(dt)
            masses.append(orc.mass())
            cfls.append(orc.cfl_dt())
    hg, gg = ctx.fields()
    ho, go = orc.fields()
    assert rel_l2(hg, ho) <= 1e-6
    assert rel_l2(hg - h0, ho - h0) <= 1e-5, rel_l2(hg - h0, ho - h0)
    assert np.max(np.abs(gg - go)) <= 1e-6
    np.testing.assert_array_equal(ctx.rupture_mask(), orc.rupture_mask())
    assert ctx.current_time() == orc.current_time()
    assert orc.active_tfe_count() > 0
    assert ctx.active_tfe_count() == orc.active_tfe_count()
    np.testing.assert_array_equal(ctx.tfe_mask(), orc.tfe_mask())
    m, c = ctx.history(60)
    np.testing.assert_allclose(m, masses, rtol=1e-9)
    np.testing.assert_allclose(c, cfls, rtol=1e-9)


def test_two_step_spectral_frames(monkeypatch):
    """sf_drainage_step_spectral_frames with two steps per launch (both
    states of each launch rendered, the first from the state the launch
    writes between its steps) against one step per launch: every frame
    bit-identical, odd step counts and fronts included."""
    import ctypes as C

    monkeypatch.setenv("SOAPFILM_STENCIL", "vec4")
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    nt, npf = 66, 768
    n = nt * npf
    p = params(HEADLINE_FLAGS)
    h0 = W.smooth_bubble(nt, npf, seed=6)
    g0 = np.full_like(h0, 0.5)
    cam = W.camera(p.bubble_radius)
    camera = L.Camera()
    for k in range(3):
        camera.position[k] = cam[k]
    cmf = W.spectral_cmf(3).astype(np.float32)
    frames = {}
    for spl in (1, 2):
        ctx = DrainageContext(p, nt, npf, 500e-9)
        ctx.set_steps_per_launch(spl)
        assert ctx.steps_per_launch() == spl
        ctx.set_fields(h0, g0)
        dt = ctx.stats().dt_cfl * p.cfl_safety
        out = []
        for count in (7, 24, 9):          # fronts spawn after 0.5 s (about step 12)
            buf = C.c_void_p()
            size = 3 * n * count * 4
            assert hip.hipMalloc(C.byref(buf), size) == 0
            try:
                L.check(L.lib().sf_drainage_step_spectral_frames(ctx.handle, count, dt, C.byref(camera), 1.33,
                                                                  L.fptr(cmf), buf))
                L.check(L.lib().sf_drainage_synchronize(ctx.handle))
                host = np.empty((count, 3, n), np.float32)
                assert hip.hipMemcpy(host.ctypes.data, buf, size, 2) == 0
            finally:
                hip.hipFree(buf)
            out.append(host)
        assert ctx.active_tfe_count() > 0
        frames[spl] = np.concatenate(out)
        ctx.close()
    assert frames[1].shape == (40, 3, n)
    np.testing.assert_array_equal(frames[2], frames[1])
    # consecutive frames differ: each step's state is rendered
    assert all(np.any(frames[1][k] != frames[1][k + 1]) for k in range(39))
