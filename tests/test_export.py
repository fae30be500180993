"""Frame export: the thickness field rendered through the runtime LUT
(bubble.wgsl:425-439, sampler of pipeline.rs:350-359) and PNG writing
(image_export.rs:42-80, with its unit tests 85-109 restated)."""
import numpy as np
import pytest

import oracle as O
from soap_bubble_simulation_amd.export import ExportError, export_frame, read_png_rgba


def test_export_frame_red_2x2(tmp_path):                      # image_export.rs:85-98
    data = np.tile(np.array([255, 0, 0, 255], np.uint8), 4)
    path = str(tmp_path / "t.png")
    export_frame(path, 2, 2, data)
    assert np.array_equal(read_png_rgba(path).reshape(-1), data)
    PIL = pytest.importorskip("PIL.Image")                     # independent decoder
    img = np.asarray(PIL.open(path).convert("RGBA"))
    assert np.array_equal(img.reshape(-1), data)


def test_export_invalid_dimensions(tmp_path):                  # image_export.rs:100-104
    with pytest.raises(ExportError) as e:
        export_frame(str(tmp_path / "t.png"), 0, 100, b"")
    assert e.value.kind == "InvalidDimensions"


def test_export_wrong_length(tmp_path):                        # image_export.rs:106-109
    with pytest.raises(ExportError) as e:
        export_frame(str(tmp_path / "t.png"), 10, 10, bytes(100))
    assert e.value.kind == "BufferCreation"


def test_oracle_render_closed_forms():
    """Texel centres return texels exactly; intensity scales before the clamp."""
    lut = np.random.default_rng(0).integers(0, 256, (64, 256, 4), dtype=np.uint8)
    lut[..., 3] = 255
    # thickness at texel centre i: t = (i + 0.5)/256 * 2000 nm; view cos at the
    # equator row facing the camera on the +z axis is 1 -> v clamps to the last row
    nt, npf = 3, 4
    i = 37
    h = np.full((nt, npf), (i + 0.5) / 256 * 2000e-9)
    cam = (0.0, 0.0, 1.0)
    out = O.render_lut(h, 0.01, cam, lut, intensity=1.0)
    # cell (1, 1): theta = phi = pi/2, normal (0, 0, 1) faces the camera -> cos 1
    # -> v = 1 -> y = 63.5 -> rows 63 and 64 clamped to 63
    assert np.array_equal(out[1, 1, :3], lut[63, i, :3])
    out2 = O.render_lut(h, 0.01, cam, lut, intensity=2.0)
    want = np.minimum(lut[63, i, :3].astype(np.float64) / 255 * 2, 1.0)
    assert np.array_equal(out2[1, 1, :3], np.floor(want * 255 + 0.5).astype(np.uint8))


@pytest.mark.gpu
def test_gpu_render_matches_oracle(tmp_path):
    from soap_bubble_simulation_amd import _lib as L
    from soap_bubble_simulation_amd import workloads as W
    from soap_bubble_simulation_amd.drainage import DrainageContext
    from soap_bubble_simulation_amd.export import render_frame
    from soap_bubble_simulation_amd.interference import interference_lut_rgba8
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    nt, npf = 96, 192
    ctx = DrainageContext(p, nt, npf, 500e-9)
    h0 = W.smooth_bubble(nt, npf, seed=2)
    ctx.set_fields(h0, np.full_like(h0, 0.5))
    cam = W.camera(p.bubble_radius)
    frame = render_frame(ctx, camera=cam, refractive_index=1.33, intensity=4.0)
    lut = np.asarray(interference_lut_rgba8(1.33, 1.0), np.uint8).reshape(64, 256, 4)
    h, _ = ctx.fields()
    ref = O.render_lut(h, p.bubble_radius, cam, lut, intensity=4.0)
    diff = np.abs(frame.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 0.01
    assert np.all(frame[..., 3] == 255)
    path = str(tmp_path / "bubble.png")
    export_frame(path, npf, nt, frame)
    assert np.array_equal(read_png_rgba(path), frame)


def test_export_refuses_non_byte_data(tmp_path):
    """The reference takes &[u8]: wider or float data is refused, not cast."""
    with pytest.raises(TypeError):
        export_frame(str(tmp_path / "t.png"), 1, 1, np.array([1.0, 0.0, 0.0, 1.0]))
    with pytest.raises(TypeError):
        export_frame(str(tmp_path / "t.png"), 1, 1, np.array([255, 0, 0, 255], np.int32))


def test_render_frame_refuses_wrong_lut():
    """render_frame checks the user LUT before the C call reads 65536 bytes."""
    from soap_bubble_simulation_amd.export import render_frame

    class Ctx:                                 # never reaches the library
        n_theta, n_phi, handle = 4, 8, None
    for bad in (np.zeros((64, 256, 3), np.uint8), np.zeros((1024, 90, 3), np.uint8),
                np.zeros((64, 256, 4), np.float32), np.zeros(100, np.uint8)):
        with pytest.raises(ValueError):
            render_frame(Ctx(), lut=bad)
