"""Two drainage steps per foam launch (foam_step2_kernel) against one-step launches.

A two-step launch streams every bubble's state once for two steps, re-forming
its band's halo rows, its column chunk's halo columns and, for bands at a
pole, the pole's ring row of the first step over every chunk; every cell uses
the one-step kernel's arithmetic, so the state after any number of steps must
equal that of one-step launches bit for bit (compared as the FP64 value of the
float-float pair, pole rows included, and Γ).  The FP64 oracle check of the
one-step path (test_foam_gpu.py) then carries over.
"""
import numpy as np
import pytest

from soap_bubble_simulation_amd import _lib as L
from soap_bubble_simulation_amd.foam import FoamBubble, FoamDrainage, foam_params

from test_foam_gpu import make_foam, oracle_for, rel_l2, seeded_fields

pytestmark = pytest.mark.gpu

# Every code path and halo case.  Where two steps are possible a chunk holds
# 62 lanes' columns (lanes 0 and 63 form its halo columns): n_phi = 2 mod 4
# (two-float2 lanes: 10, 66, 130, 250, 258), a last chunk of exactly one lane
# (126 at 2 columns per lane, 252 at 4), a lone two-column tail lane (250 at
# 4 columns per lane: lane 1 holds the chunk, the west halo is lane 0 and the
# east halo that lane's third column), single-chunk bubbles whose halo columns
# wrap onto their own edge columns (8, 10, 40, 66 at 2 or 4 columns per lane),
# several chunks (258, 520), several bands (98, 128, 248 interior rows) and a
# 4-row bubble (two interior rows).
SHAPES2 = [(4, 8), (9, 10), (20, 40), (33, 66), (65, 130), (100, 200), (64, 128), (130, 260),
           (100, 258), (260, 520), (40, 126), (70, 252), (50, 250)]


def pair(shapes, seed, shape_idx=None):
    """Two foams with identical bubbles and fields, one and two steps per launch."""
    fa, p = make_foam(shapes, seed=seed)
    fb, _ = make_foam(shapes, seed=seed)
    if shape_idx is not None:
        fa.set_tile_shapes([shape_idx] * 4)
        fb.set_tile_shapes([shape_idx] * 4)
    fa.set_steps_per_launch(1)
    fb.set_steps_per_launch(2)
    assert fa.steps_per_launch() == 1 and fb.steps_per_launch() == 2
    h, g = seeded_fields(fa, seed=seed + 100)
    fa.set_fields(h, g)
    fb.set_fields(h, g)
    return fa, fb, p, h, g


def assert_same(fa, fb, what):
    ha, ga = fa.fields()
    hb, gb = fb.fields()
    bad = np.flatnonzero(ha != hb)
    assert bad.size == 0, "%s: h differs at %d cells, first %s (%r vs %r)" % (
        what, bad.size, bad[:5], ha[bad[:3]], hb[bad[:3]])
    assert np.array_equal(ga, gb), what
    sa, sb = fa.stats(), fb.stats()
    for k in ("mass", "dt_cfl", "min", "max"):
        assert np.array_equal(sa[k], sb[k]), (what, k)


@pytest.mark.parametrize("shape_idx", range(10))
def test_foam_two_step_bit_identical(shape_idx):
    fa, fb, _, _, _ = pair(SHAPES2, seed=40 + shape_idx, shape_idx=shape_idx)
    # 37 steps: the first step reads the stored pole rows (one step), the next
    # re-aligns to an even renormalisation phase, then pairs cross several
    # rebases and a 16-step graph replay, and the last step is a single one.
    for n in (1, 37, 2, 16):
        fa.step(n)
        fb.step(n)
        assert_same(fa, fb, "shape %d after %d more steps" % (shape_idx, n))


def test_foam_two_step_without_graphs_and_timed():
    fa, fb, _, _, _ = pair(SHAPES2, seed=61)
    fb.use_graph(False)
    fb.kernel_timing(True)
    fa.step(20)
    fb.step(20)
    ms, n = fb.kernel_timing(False)
    assert ms > 0 and n > 0
    assert_same(fa, fb, "direct two-step launches")


def test_foam_two_step_active_mask():
    fa, fb, _, _, _ = pair(SHAPES2, seed=62)
    mask = np.array([1, 0, 1, 1, 0, 1, 0, 1, 1, 0, 1, 0, 1], np.uint8)
    for f in (fa, fb):
        f.step(5)
        f.set_active(mask)
        f.step(12)
        f.set_active(None)
        f.step(6)
    assert_same(fa, fb, "with held bubbles")


def test_foam_two_step_switching():
    """Switching the launch mode between calls keeps the trajectory."""
    fa, fb, _, _, _ = pair(SHAPES2, seed=63)
    fa.step(24)
    fb.step(7)
    fb.set_steps_per_launch(1)
    fb.step(5)
    fb.set_steps_per_launch(2)
    fb.step(12)
    assert_same(fa, fb, "after switching")


def test_foam_two_step_matches_oracle(oracle):
    fa, fb, p, h, g = pair(SHAPES2, seed=64)
    dts = fb.bubble_dt()
    n_steps = 41
    fb.step(n_steps)
    hg, gg = fb.fields()
    for k, b in enumerate(fb.bubbles):
        orc = oracle_for(oracle, p, b, dts[k])
        h0 = fb.split(h)[k]
        orc.set_fields(h0, fb.split(g)[k])
        orc.step(dts[k], n_steps)
        ho, go = orc.fields()
        hk, gk = fb.split(hg)[k], fb.split(gg)[k]
        err, inc = rel_l2(hk, ho), rel_l2(hk - h0, ho - h0)
        print("two-step foam bubble %d (%dx%d): rel-L2 h %.3e, increment %.3e" % (k, b.n_theta, b.n_phi, err, inc))
        assert err <= 1e-6 and inc <= 1e-5, (k, err, inc)
        assert np.max(np.abs(gk - go)) <= 1e-6 * n_steps, k


def test_foam_two_step_eligibility():
    fm, _ = make_foam([(3, 4), (20, 40)], seed=65)
    assert fm.steps_per_launch() == 1
    with pytest.raises(L.SoapFilmError):
        fm.set_steps_per_launch(2)
    with pytest.raises(L.SoapFilmError):
        fm.set_steps_per_launch(3)
    small, _ = make_foam([(20, 40)], seed=66)
    assert small.steps_per_launch() == 1            # latency-bound size: one step by default
    small.set_steps_per_launch(2)
    assert small.steps_per_launch() == 2
