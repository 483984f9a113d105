"""Physical constants and the quantity type (src/scopyon/constants.py:1-17).

``N_A`` is the exact SI Avogadro number; ``hc`` is the product of the exact SI
Planck constant and speed of light, formed in float64.
"""
from .units import Q_, Quantity, ureg  # noqa: F401

__all__ = ['ureg', 'Quantity', 'Q_', 'AVOGADROS_NUMBER', 'N_A', 'hc']

AVOGADROS_NUMBER = 6.02214076e23
N_A = AVOGADROS_NUMBER
hc = 6.62607015e-34 * 299792458.0
