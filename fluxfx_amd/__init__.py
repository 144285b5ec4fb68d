"""fluxfx_amd: MI355X-native dense smoke/fire step.

Python host classes with the call surface of FluxFX's backend (mac, pressure,
multigrid, combustion, projected) over hand-written gfx950 HIP kernels in
libfluxhip.so (C-ABI, see include/fluxhip.h).
"""
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libfluxhip.so"
__version__ = "0.1.0"
