"""mrmustard_amd: MI355X-native Bargmann->Fock amplitudes for MrMustard.

``from mrmustard_amd import math`` gives the backend manager whose
``hermite_renormalized(A, b, c, shape, stable=False, out=None)`` runs the HIP
kernels of ``libmmhip.so`` (see ``include/mm_hermite.h``).
"""

from .math import math
from .settings import settings

__version__ = "0.1.0"

__all__ = ["math", "settings"]
