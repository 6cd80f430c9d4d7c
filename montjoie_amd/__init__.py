"""MI355X-native explicit spectral-element wave solver (Montjoie's leapfrog
path): HIP kernels for gfx950 behind the C ABI of include/mjwave.h."""
from ._lib import MjwError, lib  # noqa: F401
from .problem import (TEMPORAL_SOURCES, Mesh, WaveProblem, layered_values,  # noqa: F401
                      regular_mesh, ricker_pulse_table, temporal_source_table,
                      time_source_table)

__all__ = ["MjwError", "lib", "Mesh", "WaveProblem", "regular_mesh", "layered_values",
           "ricker_pulse_table", "time_source_table", "temporal_source_table",
           "TEMPORAL_SOURCES"]
