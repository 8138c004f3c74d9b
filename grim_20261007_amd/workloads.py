"""Geometry-snapshot builders for the benchmark and test workloads.

Every builder returns a snapshot in the reference's input layout
(``{"segments": [...], "ibcs": [], "dielectrics": []}``, one TYPE 2 PEC
segment whose primitives are single panels, ``properties = ["2", "1", "0",
"0", "0"]``), drawn clockwise so that the user-facing normal points into the
air, as ``tools/GHOST/ghost_backend/twod/geometry.py:1009-1065`` asks for a
TYPE 2 contour.  Coordinates are in meters.

Workloads (sizes as used by the bench and the parity tests):

* ``circle``: regular N-gon on a circle, vertices at angle -2 pi i / N
  (the reference fixture convention, ``tests/test_rcs_physics_regression.py:50-67``).
* ``naca0012``: NACA 0012 section, chord c, closed trailing edge
  (thickness coefficient -0.1036), cosine-spaced stations, ``n_side``
  panels per surface, so N = 2 n_side.
* ``star``: r(t) = R (1 + sum_k a_k cos(k t + phi_k)), optional x/y aspect
  stretch, N panels uniform in the curve parameter.
"""
from __future__ import annotations

import math

import numpy as np

C0 = 299_792_458.0


def _snapshot(points: np.ndarray, name: str) -> dict:
    """Closed chain of points (last point not repeated) -> one TYPE 2 segment."""
    pts = np.asarray(points, dtype=float)
    n = len(pts)
    pairs = []
    for i in range(n):
        a, b = pts[i], pts[(i + 1) % n]
        pairs.append({"x1": float(a[0]), "y1": float(a[1]),
                      "x2": float(b[0]), "y2": float(b[1])})
    return {"segments": [{"name": name, "seg_type": 2,
                          "properties": ["2", "1", "0", "0", "0"],
                          "point_pairs": pairs}],
            "ibcs": [], "dielectrics": []}


def signed_area(points: np.ndarray) -> float:
    p = np.asarray(points, dtype=float)
    q = np.roll(p, -1, axis=0)
    return 0.5 * float(np.sum(p[:, 0] * q[:, 1] - q[:, 0] * p[:, 1]))


def circle_points(radius: float, count: int) -> np.ndarray:
    theta = np.linspace(0.0, -2.0 * np.pi, count + 1)[:-1]
    return np.stack((radius * np.cos(theta), radius * np.sin(theta)), axis=1)


def circle(radius: float = 1.0, count: int = 400) -> dict:
    return _snapshot(circle_points(radius, count), "circle")


def naca0012_points(n_side: int, chord: float = 1.0) -> np.ndarray:
    """Clockwise contour: trailing edge -> lower surface -> leading edge -> upper."""
    beta = np.linspace(0.0, np.pi, n_side + 1)
    x = 0.5 * (1.0 - np.cos(beta))            # 0 (LE) .. 1 (TE), cosine spaced
    t = 0.12
    yt = 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2
                    + 0.2843 * x**3 - 0.1036 * x**4)
    yt[0] = 0.0
    yt[-1] = 0.0
    # lower surface from TE (x=1) to LE (x=0), then upper from LE to TE (exclusive).
    lower = np.stack((x[::-1], -yt[::-1]), axis=1)          # TE .. LE
    upper = np.stack((x[1:-1], yt[1:-1]), axis=1)           # after LE .. before TE
    pts = np.concatenate((lower, upper), axis=0) * chord
    return pts


def naca0012(n_side: int = 1000, chord: float = 1.0) -> dict:
    return _snapshot(naca0012_points(n_side, chord), "naca0012")


def star_points(count: int, radius: float = 0.5, coefficients=None, phases=None,
                aspect: float = 1.0) -> np.ndarray:
    a = np.zeros(8) if coefficients is None else np.asarray(coefficients, float)
    phi = np.zeros(len(a)) if phases is None else np.asarray(phases, float)
    theta = np.linspace(0.0, -2.0 * np.pi, count + 1)[:-1]
    r = np.ones_like(theta)
    for k in range(len(a)):
        r = r + a[k] * np.cos((k + 1) * theta + phi[k])
    r = radius * r
    sx, sy = math.sqrt(aspect), 1.0 / math.sqrt(aspect)
    return np.stack((sx * r * np.cos(theta), sy * r * np.sin(theta)), axis=1)


def star_shape(rng: np.random.Generator, amplitude: float = 0.1):
    """Coefficients a_k ~ U(0, amplitude) and phases phi_k ~ U(0, 2 pi), k = 1..8."""
    return rng.uniform(0.0, amplitude, 8), rng.uniform(0.0, 2.0 * np.pi, 8)


def star(count: int, radius: float = 0.5, coefficients=None, phases=None,
         aspect: float = 1.0) -> dict:
    return _snapshot(star_points(count, radius, coefficients, phases, aspect), "star")


def contour_perimeter(points: np.ndarray) -> float:
    p = np.asarray(points, float)
    return float(np.sum(np.linalg.norm(np.roll(p, -1, axis=0) - p, axis=1)))


def frequency_for_ka(ka: float, radius: float) -> float:
    """Frequency in GHz with k * radius = ka."""
    return ka / radius * C0 / (2.0 * math.pi) / 1e9


def frequency_for_segments_per_wavelength(points: np.ndarray, per_wavelength: float) -> float:
    """Frequency (GHz) at which the LONGEST panel is 1/per_wavelength of a wavelength."""
    p = np.asarray(points, float)
    longest = float(np.max(np.linalg.norm(np.roll(p, -1, axis=0) - p, axis=1)))
    wavelength = longest * per_wavelength
    return C0 / wavelength / 1e9


def star_sweep(seed: int = 20261203, nfreq: int = 64, fmin_ghz: float = 0.3, fmax_ghz: float = 3.0,
               per_wavelength: float = 100.0, radius: float = 0.5, nangles: int = 360):
    """Batched frequency sweep of one random star-shaped PEC curve.

    N follows the mesh density: ``per_wavelength`` panels per wavelength of the
    LONGEST panel at each frequency.  At 10 panels per wavelength the
    R = 0.5 m curves give N of about 35-370; the sweep's stated range of about
    100-3200 unknowns needs about 100 panels per wavelength, the default here.
    Returns a list of (points, frequency_ghz, angles).
    """
    rng = np.random.default_rng(seed)
    a, phi = star_shape(rng)
    freqs = np.logspace(np.log10(fmin_ghz), np.log10(fmax_ghz), nfreq)
    probe = star_points(4096, radius, a, phi)
    perim = contour_perimeter(probe)
    out = []
    for f in freqs:
        lam = C0 / (f * 1e9)
        n = max(32, int(math.ceil(perim / lam * per_wavelength)))
        out.append((star_points(n, radius, a, phi), float(f), np.arange(nangles) * (360.0 / nangles)))
    return out


def serving_stream(count: int, seed: int = 7, nmin: int = 200, nmax: int = 12000, reuse: float = 0.3,
                   angle_sets=(90, 360, 1440), per_wavelength: float = 10.0, radius: float = 0.5):
    """Requests of the serving configuration: star-shaped curves, N log-uniform in [nmin, nmax],
    angle-set size drawn from angle_sets, a fraction ``reuse`` reusing an earlier (geometry,
    frequency) with a new angle set.  Frequency is set so the longest panel is
    lambda / per_wavelength.  Returns a list of dicts (points, frequency_ghz, angles, key)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nangles = int(rng.choice(angle_sets))
        angles = np.sort(rng.uniform(0.0, 360.0, nangles))
        if out and rng.uniform() < reuse:
            base = out[int(rng.integers(len(out)))]
            out.append({"points": base["points"], "frequency_ghz": base["frequency_ghz"], "angles": angles,
                        "key": base["key"], "reused": True})
            continue
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        a, phi = star_shape(rng)
        pts = star_points(n, radius, a, phi)
        f = frequency_for_segments_per_wavelength(pts, per_wavelength)
        out.append({"points": pts, "frequency_ghz": f, "angles": angles, "key": f"req{i}", "reused": False})
    return out
