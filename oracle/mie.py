"""CPU ORACLE -- test infrastructure only, never part of the product path.

Closed-form Mie (cylindrical-wave) series for the monostatic 2-D scattering of
a PEC circular cylinder, in the solver's amplitude convention: the stored
amplitude B satisfies sigma_2D = |B|^2 / (4 k) and B = -4j sum_n a_n (-1)^n
with a_n = -J_n(ka) / H_n^(2)(ka) (TM, Dirichlet) or
a_n = -J_n'(ka) / H_n^(2)'(ka) (TE, Neumann), n from -M to M,
M = ceil(ka + 4.05 (ka)^(1/3) + 2) + 10 (Wiscombe's rule plus a pad).
Follows tools/GHOST/ghost_backend/validation/cylinder.py:30-90.

Pinned in tests/test_oracle_mie.py, without running the reference's code:
* the small-ka (Rayleigh) limits derived independently of the series' sum:
  TM B -> -4j a_0 with a_0 = -1 / (1 - (2j/pi)(ln(ka/2) + gamma_E)), and
  TE B -> 3 pi (ka)^2 (the n = 0 and n = +-1 terms, so a wrong sign of the
  (-1)^n factor or of the n = +-1 coefficients fails);
* losslessness of every PEC coefficient, |1 + 2 a_n| = 1;
* the reference's own cylinder gates (tests/test_rcs_physics_regression.py:
  648-669 and 671-719) re-run with the oracle's BEM solve, itself pinned to
  the reference's golden vectors: the 96-chord ka = 1 circle within 0.01 dB
  in both polarizations, and the complex amplitude of the 64-chord circle at
  the first zero of J_0 within 5e-3;
* the reference-written golden amplitudes of the 64-chord ka = 1 and the
  400-chord ka = 10 circles (tests/golden), which agree with the series to
  the discretisation error (1.7e-3 and 4.3e-4 relative).

``dielectric_backscatter_amplitude`` is the homogeneous dielectric cylinder's
series (validation/cylinder.py:133-191): the interior field J_n(k1 r) matched to
the exterior J_n + a_n H_n^(2) by continuity of u and of (1/zeta) du/dr, zeta =
mu_r (TM) or eps_r (TE).  Pinned in tests/test_dielectric_cpu.py: eps = mu = 1
gives a_n = 0 exactly (no scatterer); eps -> infinity in TM (zeta = mu fixed)
tends to the PEC Dirichlet series; lossless media give |1 + 2 a_n| = 1; and the
reference-written dielectric golden amplitudes agree with the series to the
discretisation error.

``impedance_backscatter_amplitude`` (validation/cylinder.py:93-131) is the
Leontovich cylinder: each harmonic of the total field satisfies du/dr + beta u = 0
at the surface, beta = -j k eta0 / Zs (TM) or -j k Zs / eta0 (TE).
``coated_impedance_backscatter_amplitude`` (validation/cylinder.py:375-448) puts
that surface at the inner radius of an annular medium: the annulus field
J_n(k1 r) + c Y_n(k1 r) takes c from the Robin condition (with the medium's k1
and eta1 = eta0 mu / n), then matches the exterior as the dielectric series does.
Pinned in tests/test_lossy_cpu.py: Zs -> 0 tends to the PEC series in both
polarizations; a coating of eps = mu = 1 gives the bare impedance cylinder of the
inner radius; and the coated impedance core's solves converge to it at second
order (tests/test_gpu_regions.py).
"""
import math

import numpy as np
from scipy import special as sp

C0 = 299_792_458.0
ETA0 = 376.730313668


def terms_for_ka(ka: float, pad: int = 10) -> int:
    n = int(math.ceil(abs(ka) + 4.05 * abs(ka) ** (1.0 / 3.0) + 2.0))
    return max(10, n + pad)


def pec_coefficients(ka: float, pol: str):
    """(n, a_n) for n = -M..M."""
    m = terms_for_ka(ka)
    n = np.arange(-m, m + 1)
    if pol.upper() == "TM":
        a = -sp.jv(n, ka) / sp.hankel2(n, ka)
    elif pol.upper() == "TE":
        a = -sp.jvp(n, ka, 1) / sp.h2vp(n, ka, 1)
    else:
        raise ValueError(pol)
    return n, a


def pec_backscatter_amplitude(radius: float, k: float, pol: str) -> complex:
    n, a = pec_coefficients(k * radius, pol)
    return complex(-4.0j * np.sum(a * (-1.0) ** n))


def pec_sigma(radius: float, k: float, pol: str) -> float:
    return abs(pec_backscatter_amplitude(radius, k, pol)) ** 2 / (4.0 * k)


def dielectric_coefficients(ka: float, eps_r: complex, mu_r: complex, pol: str):
    """(n, a_n) of a homogeneous cylinder of relative (eps_r, mu_r) at k0 a = ka."""
    index = np.sqrt(complex(eps_r) * complex(mu_r))
    mx = index * ka
    m = terms_for_ka(max(abs(ka), abs(mx)))
    n = np.arange(-m, m + 1)
    zeta = complex(mu_r) if pol.upper() == "TM" else complex(eps_r)
    if pol.upper() not in ("TM", "TE"):
        raise ValueError(pol)
    p, q = index / zeta, 1.0           # k1 / zeta and k0 / 1, both divided by k0
    num = -(p * sp.jvp(n, mx, 1) * sp.jv(n, ka) - q * sp.jv(n, mx) * sp.jvp(n, ka, 1))
    den = p * sp.jvp(n, mx, 1) * sp.hankel2(n, ka) - q * sp.jv(n, mx) * sp.h2vp(n, ka, 1)
    return n, num / den


def dielectric_backscatter_amplitude(radius: float, k: float, eps_r: complex, mu_r: complex, pol: str) -> complex:
    n, a = dielectric_coefficients(k * radius, eps_r, mu_r, pol)
    return complex(-4.0j * np.sum(a * (-1.0) ** n))


def impedance_backscatter_amplitude(radius: float, k: float, zs: complex, pol: str) -> complex:
    """Leontovich cylinder of surface impedance zs in air."""
    ka = k * radius
    m = terms_for_ka(ka)
    n = np.arange(-m, m + 1)
    if pol.upper() == "TM":
        beta = -1j * k * ETA0 / complex(zs)
    elif pol.upper() == "TE":
        beta = -1j * k * complex(zs) / ETA0
    else:
        raise ValueError(pol)
    a = -(k * sp.jvp(n, ka, 1) + beta * sp.jv(n, ka)) / (k * sp.h2vp(n, ka, 1) + beta * sp.hankel2(n, ka))
    return complex(-4.0j * np.sum(a * (-1.0) ** n))


def coated_impedance_backscatter_amplitude(a_in: float, a_out: float, k0: float, eps_r: complex, mu_r: complex,
                                           zs: complex, pol: str) -> complex:
    """Impedance surface of radius a_in under an annulus of (eps_r, mu_r) out to a_out, in air."""
    index = np.sqrt(complex(eps_r) * complex(mu_r))
    k1 = k0 * index
    eta1 = ETA0 * complex(mu_r) / index
    if pol.upper() == "TM":
        zeta, beta = complex(mu_r), -1j * k1 * eta1 / complex(zs)
    elif pol.upper() == "TE":
        zeta, beta = complex(eps_r), -1j * k1 * complex(zs) / eta1
    else:
        raise ValueError(pol)
    xi, xo1, xo = k1 * a_in, k1 * a_out, k0 * a_out
    m = terms_for_ka(max(abs(xi), abs(xo1), abs(xo)))
    n = np.arange(-m, m + 1)
    c = -(k1 * sp.jvp(n, xi, 1) + beta * sp.jv(n, xi)) / (k1 * sp.yvp(n, xi, 1) + beta * sp.yv(n, xi))
    phi = sp.jv(n, xo1) + c * sp.yv(n, xo1)
    dphi = sp.jvp(n, xo1, 1) + c * sp.yvp(n, xo1, 1)
    p, q = k1 / zeta, k0
    a = -(p * dphi * sp.jv(n, xo) - q * phi * sp.jvp(n, xo, 1)) / (p * dphi * sp.hankel2(n, xo) -
                                                                    q * phi * sp.h2vp(n, xo, 1))
    return complex(-4.0j * np.sum(a * (-1.0) ** n))
