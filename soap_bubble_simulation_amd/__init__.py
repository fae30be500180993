"""MI355X-native (gfx950) soap-film physics: drainage step and thin-film
interference of pjt222/soap_bubble_simulation, as hand-written HIP kernels
behind the C ABI of ``include/soapfilm.h``.

Python modules mirror the reference's Rust API for this path:

* ``drainage``      DrainageSimulator, ThicknessField, GPUDrainageSimulator,
                    DrainageParams (src/physics/drainage.rs,
                    src/render/gpu_drainage.rs)
* ``interference``  InterferenceCalculator, generate_interference_lut
                    (src/physics/interference.rs, src/render/interference_lut.rs)
* ``config``        SimulationConfig and its JSON form (src/config.rs)
* ``workloads``     seeded synthetic inputs of the tests and the benchmark
"""
from ._lib import SoapFilmError, device_available  # noqa: F401

__all__ = ["SoapFilmError", "device_available"]
__version__ = "0.1.0"
