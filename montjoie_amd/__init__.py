"""MI355X-native explicit spectral-element wave solver (Montjoie's leapfrog
path): HIP kernels for gfx950 behind the C ABI of include/mjwave.h."""
from ._lib import MjwError, lib  # noqa: F401
from .problem import (Mesh, WaveProblem, layered_values, regular_mesh,  # noqa: F401
                      ricker_pulse_table, time_source_table)

__all__ = ["MjwError", "lib", "Mesh", "WaveProblem", "regular_mesh", "layered_values",
           "ricker_pulse_table", "time_source_table"]
