"""Two drainage steps per stencil launch (temporal blocking).

``drainage_band2_kernel`` streams a band once per two steps.  Its cell
update is the one-step kernels' (drainage.rs:298-392 interior step with the
surfactant transport of drainage.wgsl:178-201), and the pole means between
its two steps are summed in the one-step order, so the bar is equality bit
for bit with one step per launch: every (hi, lo) pair, Γ, the pole rows and
the simulation time.  Both contexts take the float4 kernels
(SOAPFILM_STENCIL=vec4) so that their column tiles, and with them the pole
partial sums, are the same.  One trajectory is also compared with the FP64
oracle directly, at the bars of the one-step parity tests.
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
    # the comparison has teeth: the film and, when transported, Γ moved
    assert rel_l2(h1 - hs, hs) > 1e-9
    if flags & L.FLAG_SURFACTANT:
        assert np.max(np.abs(g1 - gs)) > 1e-6
    return one, two, h0


@pytest.mark.parametrize("flags", [L.FLAG_SURFACTANT, 0, L.FLAG_SURFACTANT | L.FLAG_POLES_COPY_H,
                                   L.FLAG_POLES_COPY_H])
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


def test_two_step_ineligible_flags_keep_one_step(monkeypatch):
    monkeypatch.setenv("SOAPFILM_STENCIL", "vec4")
    for flags in (L.FLAG_TFE, L.FLAG_DIAG, L.FLAG_MARANGONI,
                  L.FLAG_SURFACTANT | L.FLAG_POLES_AVG_GAMMA):
        c = DrainageContext(params(flags), 66, 512, 500e-9)
        c.set_steps_per_launch(2)
        assert c.steps_per_launch() == 1, flags
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
