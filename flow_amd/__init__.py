"""flow_amd -- MI355X-native hip kernel set for UoB-HPC/flow's hydro timestep.

The product is the C-ABI library ``flow_amd/lib/libflowhip.so`` (HIP kernels
for gfx950 behind flow_interface.h, plus the arch-compatible layer). This
package is a thin ctypes host over it for tests and benchmarks; it never
computes the timestep itself, and importing the runtime fails loudly when the
library has not been built.
"""
from .native import (  # noqa: F401
    ARRAY_NAMES, CFL_DEFINED, CFL_OMP3_SERIAL, FUSED, MARCH, STAGEWISE, Mesh,
    MeshStruct, DeviceState,
    DeviceState32, Ensemble, lib,
    library_path,
)
from . import problems  # noqa: F401
