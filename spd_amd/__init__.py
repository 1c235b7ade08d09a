"""spd_amd: MI355X-native SD-ADER time step for the compressible Euler equations.

The public entry point is :class:`spd_amd.simulator.SDADER_Simulator`, which
keeps the constructor keywords, methods and snapshot attributes of the
reference simulator (``src/sdader_simulator.py:14``).  The arithmetic of the
step runs in hand-written HIP kernels for gfx950 behind the C ABI declared in
``include/spd.h`` (library ``spd_amd/libspd_hip.so``).

Importing the package does not load the GPU library; it is loaded on first
use and a missing library raises ``RuntimeError``.
"""
from . import initial_conditions  # noqa: F401  (pure NumPy)

__all__ = ["SDADER_Simulator", "initial_conditions"]


def __getattr__(name):
    if name == "SDADER_Simulator":
        from .simulator import SDADER_Simulator
        return SDADER_Simulator
    raise AttributeError(name)
