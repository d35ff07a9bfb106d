"""dftpy_amd: MI355X-native orbital-free DFT energy/potential evaluation,
density optimizer and Crank-Nicolson propagator with DFTpy's Python
interface (shaoxc/dftpy), computing through libofdft_hip (HIP, gfx950)."""
from . import _lib  # noqa: F401  (fails loudly when the library is not built)
from .field import DirectField, ReciprocalField
from .grid import DirectGrid, ReciprocalGrid
from .ions import Ions

__version__ = "0.1.0"
__all__ = ["DirectGrid", "ReciprocalGrid", "DirectField", "ReciprocalField", "Ions"]
