"""CPU ORACLE -- test infrastructure only.

Closed-form Mie (cylindrical-wave) series for the monostatic 2-D scattering of
a PEC circular cylinder, in the solver's amplitude convention: the stored
amplitude B satisfies sigma_2D = |B|^2 / (4 k) and B = -4j sum_n a_n (-1)^n
with a_n = -J_n(ka) / H_n^(2)(ka) (TM, Dirichlet) or
a_n = -J_n'(ka) / H_n^(2)'(ka) (TE, Neumann), n from -M to M,
M = ceil(ka + 4.05 (ka)^(1/3) + 2) + 10 (Wiscombe's rule plus a pad).
Follows tools/GHOST/ghost_backend/validation/cylinder.py:30-90.
Pinned in tests/test_oracle_mie.py by the small-ka Rayleigh limit and by the
reference solver's own cylinder fixtures.
"""
import math

import numpy as np
from scipy import special as sp

C0 = 299_792_458.0


def terms_for_ka(ka: float, pad: int = 10) -> int:
    n = int(math.ceil(abs(ka) + 4.05 * abs(ka) ** (1.0 / 3.0) + 2.0))
    return max(10, n + pad)


def pec_backscatter_amplitude(radius: float, k: float, pol: str) -> complex:
    ka = k * radius
    m = terms_for_ka(ka)
    n = np.arange(-m, m + 1)
    if pol.upper() == "TM":
        a = -sp.jv(n, ka) / sp.hankel2(n, ka)
    elif pol.upper() == "TE":
        a = -sp.jvp(n, ka, 1) / sp.h2vp(n, ka, 1)
    else:
        raise ValueError(pol)
    return complex(-4.0j * np.sum(a * (-1.0) ** n))


def pec_sigma(radius: float, k: float, pol: str) -> float:
    return abs(pec_backscatter_amplitude(radius, k, pol)) ** 2 / (4.0 * k)
