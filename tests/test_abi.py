"""C ABI checks that need no GPU: the library loads, exports every entry
point declared in include/soapfilm.h, the ctypes mirrors match the C struct
layouts, defaults match the reference, and compute entry points refuse to
run (SF_ERR_DEVICE) when no GPU is present instead of falling back to CPU.
One GPU test runs the C++ API test binary (tests/cpp/test_api.cpp)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from soap_bubble_simulation_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "soapfilm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sf_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and the ctypes binding declares every one of them
    assert not [n for n in names if n not in L.SIGNATURES]


def test_struct_layouts():
    assert C.sizeof(L.Params) == 136
    assert L.Params.flags.offset == 128
    assert C.sizeof(L.Stats) == 64
    assert C.sizeof(L.Camera) == 24


def test_reference_defaults():
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    assert (p.gravity, p.viscosity, p.density) == (9.81, 0.001, 1000.0)      # config.rs:58-95
    assert p.diffusion_coeff == 1e-9 and p.bubble_radius == 0.025
    assert p.critical_thickness == 10.0 * 1e-9                               # drainage.rs:278
    assert p.tfe_thickness_ratio == 0.85 and p.tfe_first_spawn == 0.5       # drainage.rs:283-286
    assert p.tfe_spawn_interval == 2.0 and p.sin_guard == 1e-10
    assert p.flags == L.FLAG_TFE
    g = L.Params()
    L.check(L.lib().sf_params_gpu_default(g))
    assert g.dt == 0.001 and g.critical_thickness == 30e-9                  # gpu_drainage.rs:49-55
    assert g.sin_guard == 1e-4 and g.gamma_air == 0.072 and g.marangoni_coeff == 0.01
    assert g.flags == L.FLAG_SURFACTANT | L.FLAG_POLES_COPY_H


@pytest.mark.skipif(L.device_available(), reason="checks the no-GPU behaviour")
def test_no_cpu_fallback():
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    h = C.c_void_p()
    rc = L.lib().sf_drainage_create(C.byref(p), 16, 32, 5e-7, 0.5, C.byref(h))
    assert rc == L.SF_ERR_DEVICE and not h.value
    assert "no HIP device" in L.last_error()
    import numpy as np
    out = np.zeros(64 * 256 * 4, np.uint8)
    assert L.lib().sf_lut_rgba8(1.33, 1.0, L.u8ptr(out)) == L.SF_ERR_DEVICE
    with pytest.raises(L.SoapFilmError):
        from soap_bubble_simulation_amd.drainage import DrainageSimulator
        DrainageSimulator()


def test_argument_validation():
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    h = C.c_void_p()
    assert L.lib().sf_drainage_create(C.byref(p), 2, 32, 5e-7, 0.5, C.byref(h)) == L.SF_ERR_INVALID
    assert "n_theta >= 3" in L.last_error()
    bad = p.copy()
    bad.bubble_radius = -1.0
    assert L.lib().sf_drainage_create(C.byref(bad), 16, 32, 5e-7, 0.5, C.byref(h)) == L.SF_ERR_INVALID
    assert L.lib().sf_drainage_step(None, 1, 0.1) == L.SF_ERR_INVALID
    assert L.lib().sf_lut_rgba8(1.33, 1.0, None) == L.SF_ERR_INVALID
    assert L.lib().sf_spectral_set_operand(None, L.SPECTRAL_OPERAND_FP16) == L.SF_ERR_INVALID
    k = C.c_int32()
    assert L.lib().sf_spectral_get_operand(None, C.byref(k)) == L.SF_ERR_INVALID


def test_cpp_api_compiles():
    """The C++ mirror header compiles against the C header."""
    src = os.path.join(ROOT, "tests", "cpp", "test_api.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_api_runs_reference_unit_tests():
    """The C++ mirror of drainage.rs:616-720, interference.rs:612-792 and
    interference_lut.rs:165-205 (tests/cpp/test_api.cpp), built against
    libsoapfilm.so through soapfilm.hpp, runs on the GPU and prints OK."""
    from soap_bubble_simulation_amd import build as B
    if not os.path.exists(B.CPP_TEST_BIN):
        B.build_cpp_test()
    r = subprocess.run([B.CPP_TEST_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "OK"


def test_foam_rejects_unsupported_flags():
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))      # carries SF_FLAG_TFE
    b = (L.Bubble * 1)(L.Bubble(16, 32, 0.01, 1e-3, 500e-9, 0.0))
    h = C.c_void_p()
    assert L.lib().sf_foam_create(p, 1, b, C.byref(h)) == L.SF_ERR_INVALID
    assert "SF_FLAG_SURFACTANT" in L.last_error()
    p.flags = L.FLAG_SURFACTANT
    b[0].n_phi = 31
    assert L.lib().sf_foam_create(p, 1, b, C.byref(h)) in (L.SF_ERR_INVALID, L.SF_ERR_DEVICE)


def test_bubble_struct_layout():
    assert C.sizeof(L.Bubble) == 40
    assert L.Bubble.bubble_radius.offset == 8


def test_cluster_struct_layouts_and_defaults():
    assert C.sizeof(L.ClusterParams) == 44
    assert C.sizeof(L.ClusterBubble) == 44 and L.ClusterBubble.radius.offset == 16
    assert C.sizeof(L.ClusterConnection) == 40 and L.ClusterConnection.wall_curvature_radius.offset == 32
    p = L.ClusterParams()
    L.check(L.lib().sf_cluster_params_default(p))
    # FoamSimulator::new (foam_dynamics.rs:37-53)
    assert (p.surface_tension, p.density) == (C.c_float(0.025).value, 1000.0)
    assert p.viscosity == C.c_float(0.001).value and tuple(p.gravity) == (0.0, C.c_float(-9.81).value, 0.0)
    assert p.van_der_waals_strength == C.c_float(0.001).value and p.repulsion_stiffness == 5.0
    assert p.damping == C.c_float(0.8).value and p.coalescence_enabled == 0
    assert p.coalescence_threshold == 0.5


def test_cluster_argument_validation():
    p = L.ClusterParams()
    L.check(L.lib().sf_cluster_params_default(p))
    h = C.c_void_p()
    b = (L.ClusterBubble * 1)()
    b[0].radius = -1.0
    assert L.lib().sf_cluster_create(C.byref(p), 1, b, C.byref(h)) == L.SF_ERR_INVALID
    assert "radius" in L.last_error()
    assert L.lib().sf_cluster_step(None, 0.001, 1) == L.SF_ERR_INVALID
    if not L.device_available():
        b[0].radius = 0.01
        assert L.lib().sf_cluster_create(C.byref(p), 1, b, C.byref(h)) == L.SF_ERR_DEVICE
        assert not h.value


def test_noalloc_stats_argument_checks():
    """sf_field_stats_noalloc refuses grids beyond its fixed scratch and NULL
    pointers before touching a device, and needs a GPU otherwise."""
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    fake = C.c_void_p(256)          # never dereferenced: the calls fail first
    f = L.lib().sf_field_stats_noalloc
    assert f(C.byref(p), L.NOALLOC_MAX_THETA + 1, 8, fake, None, None, fake) == L.SF_ERR_INVALID
    assert "SF_NOALLOC_MAX_THETA" in L.last_error()
    assert f(C.byref(p), 64, 128, None, None, None, fake) == L.SF_ERR_INVALID
    assert f(C.byref(p), 64, 128, fake, None, None, None) == L.SF_ERR_INVALID
    assert f(C.byref(p), 2, 128, fake, None, None, fake) == L.SF_ERR_INVALID
    if not L.device_available():
        assert f(C.byref(p), 64, 128, fake, None, None, fake) == L.SF_ERR_DEVICE
