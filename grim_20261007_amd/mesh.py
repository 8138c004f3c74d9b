"""Host geometry: reference geometry snapshot -> closed-contour linear mesh.

Follows the reference's panel builder for PEC (TYPE 2) segments
(tools/GHOST/ghost_backend/twod/geometry.py): units (:709-715), the TYPE 2
user-convention flip that reverses each chain (:1009-1065), straight-primitive
splitting with the primitive's own end vertices (:717-732), panel counts from
the N property (:906-928, explicit counts checked against the 4-panels-per-
wavelength floor, :1865-1893), panel normal (-t_y, t_x) (:1935) and node
welding on the 1e-9 m snap (:1983-2050).  The node numbering is the
reference's (first appearance, panel by panel), and every contour must close:
element e then runs from node e to node next[e].
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C0 = 299_792_458.0
EPS = 1e-12
NODE_SNAP = 1.0e-9
DEFAULT_PANELS_PER_WAVELENGTH = 20
MIN_EXPLICIT_PANELS_PER_WAVELENGTH = 4
# A wavelength-controlled primitive gets another panel only once it exceeds the
# target panel length by this fraction of a panel (twod/geometry.py:894-903).
PANEL_COUNT_CEIL_TOLERANCE = 0.05


@dataclass
class ContourMesh:
    nodes: np.ndarray        # (N, 2) welded node coordinates
    p0: np.ndarray           # (N, 2) element start points
    p1: np.ndarray           # (N, 2) element end points
    next: np.ndarray         # (N,) node at the end of element e (element e starts at node e)
    prev: np.ndarray         # (N,) element ending at node e
    contour_starts: list = field(default_factory=list)
    impedance: np.ndarray | None = None   # (N,) complex surface impedance per element (0: PEC), or None
    ref_node: np.ndarray | None = None    # (N,) the reference's node number of node p (None: p itself)
    medium: tuple | None = None           # (flag, eps, mu) of a dielectric body (TYPE 3), or None (conductor)
    regions: np.ndarray | None = None     # (N, 2) (minus, plus) region per element of a multi-region layout
    region_media: dict | None = None      # region -> (eps, mu) of a multi-region layout (0: air)
    degree: int = 1                       # Galerkin basis degree (1, 2 or 3; twod/basis.py)

    @property
    def n(self) -> int:
        return len(self.p0)

    @property
    def dofs(self) -> int:
        """Basis nodes of a conductor contour: the element ends plus degree - 1 interior nodes per
        element (twod/basis.py:47-63)."""
        return self.n * self.degree

    def basis_elements(self) -> np.ndarray:
        """(N, degree + 1) node ids per element, local order: start, end, then the interior nodes
        at 1/d .. (d-1)/d, numbered after every end node, element by element (basis.py:47-63)."""
        d = self.degree
        ends = self.elements
        if d == 1:
            return ends
        inner = self.n + (d - 1) * np.arange(self.n)[:, None] + np.arange(d - 1)[None, :]
        return np.concatenate((ends, inner), axis=1)

    def basis_nodes(self) -> np.ndarray:
        """(dofs, 2) coordinates of the basis nodes (ends first, then interior, as basis_elements)."""
        d = self.degree
        if d == 1:
            return self.p0.copy()
        t = np.arange(1, d) / d
        inner = self.p0[:, None, :] + t[None, :, None] * (self.p1 - self.p0)[:, None, :]
        return np.concatenate((self.p0, inner.reshape(-1, 2)), axis=0)

    @property
    def elements(self) -> np.ndarray:
        return np.stack((np.arange(self.n), self.next), axis=1)

    @property
    def lengths(self) -> np.ndarray:
        return np.linalg.norm(self.p1 - self.p0, axis=1)

    @property
    def normals(self) -> np.ndarray:
        seg = self.p1 - self.p0
        t = seg / self.lengths[:, None]
        return np.stack((-t[:, 1], t[:, 0]), axis=1)

    def fix_rows(self, wave: int = 64) -> np.ndarray:
        """Rows whose previous element is not the row before (wavefront heads, contour starts)."""
        rows = np.arange(self.n)
        mask = (rows % wave == 0) | (self.prev != rows - 1)
        return rows[mask].astype(np.int32)

    def far_order(self, k: float) -> int:
        """W-calibrated far rule of the paired traversal (twod/operators.py:1502-1527): 8 points for
        degree 1, 9 for degrees 2 and 3, inside the calibrated range; 16 outside it.  (The graded
        rule of a tile never exceeds the configured 8 points, so the W floor decides.)"""
        lengths = self.lengths
        lmax, lmin = float(lengths.max()), float(lengths.min())
        if 1e-4 <= abs(k) * lmax <= 3.0 and lmin >= 1e-3 * lmax:
            return 8 if self.degree == 1 else 9
        return 16

    def load_order(self, k: float) -> int:
        """Gauss order of the polynomial plane-wave moments: max(12, degree + 3 + ceil(|k| L_max))
        (twod/basis.py:89-97; taken over the whole contour, which only adds points to elements
        whose own bound is lower)."""
        order = max(12, self.degree + 3 + int(math.ceil(abs(k) * float(self.lengths.max()))))
        if order > 256:
            raise ValueError("Polynomial excitation needs a finer wavelength mesh.")
        return order


def unit_scale(units: str) -> float:
    value = (units or "").strip().lower()
    if value in {"inch", "inches", "in"}:
        return 0.0254
    if value in {"meter", "meters", "m"}:
        return 1.0
    raise ValueError(f"Unsupported geometry units '{units}'. Use inches or meters.")


def _parse_int(token, default: int = 0) -> int:
    """Nearest integer of a numeric token (twod/geometry.py:447-451)."""
    try:
        return int(round(float(token)))
    except (TypeError, ValueError):
        return default


def _parse_props(seg: dict):
    props = list(seg.get("properties", []) or [])
    vals = [_parse_int(props[i]) if i < len(props) else 0 for i in range(5)]
    seg_type = vals[0] if props else int(seg.get("seg_type", 0))
    return seg_type, vals[1], vals[2]


def _panel_count(n_prop: int, length: float, wavelength: float) -> int:
    if length <= EPS:
        return 1
    if n_prop > 0:
        return n_prop
    per = abs(n_prop) if n_prop < 0 else DEFAULT_PANELS_PER_WAVELENGTH
    target = wavelength / per
    if not (math.isfinite(target) and target > 0):
        raise ValueError("The controlling mesh wavelength must be positive and finite.")
    return max(1, int(math.ceil(length / target - PANEL_COUNT_CEIL_TOLERANCE)))


def _bit_reversed_fraction(index: int) -> float:
    """Van der Corput (base 2) value of ``index``: 0, 1/2, 1/4, 3/4, ... ; ordering a run of
    equal primitives by it picks them spread along the chain (twod/geometry.py:1569-1578)."""
    value, step = 0.0, 0.5
    while index:
        value += step * (index & 1)
        index >>= 1
        step *= 0.5
    return value


def refined_chain_counts(base_counts, lengths, factor: float) -> list:
    """Panel counts of one segment chain of the certification fine mesh
    (twod/geometry.py:1581-1640).

    The chain gets T = max(B + 1, ceil(factor B)) panels for its B base panels.  Every
    primitive keeps at least its base count; the extra panels go where panels are longest:
    a bisection finds the smallest panel-length threshold h whose counts
    max(base, ceil(length / h)) stay within T, and what is left of T is handed out one panel
    per primitive from the longest remaining panels down.  Primitives whose panels are equally
    long (the mirror halves of a symmetric body) are refined as one group, also when that
    exceeds T by up to max(2, ceil(T / 100)) panels; a larger group takes the remaining panels
    in bit-reversed index order.
    """
    base = [max(1, int(c)) for c in base_counts]
    ls = [float(v) for v in lengths]
    b_total = sum(base)
    target = max(b_total + 1, int(math.ceil(b_total * float(factor))))
    if len(base) == 1:
        return [target]

    def counts_at(h: float) -> list:
        return [max(c, int(math.ceil(v / h))) for c, v in zip(base, ls)]

    hi = max(v / c for c, v in zip(base, ls))
    if not hi > 0.0:
        out = list(base)
        out[0] += target - b_total
        return out
    lo = hi / max(2.0, float(factor)) / 2.0
    while sum(counts_at(lo)) <= target and lo > EPS:
        lo *= 0.5
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if sum(counts_at(mid)) <= target:
            hi = mid
        else:
            lo = mid
        if hi - lo <= 1.0e-12 * hi:
            break
    out = counts_at(hi)
    left = target - sum(out)
    if left <= 0:
        return out
    longest = max(v / c for c, v in zip(out, ls))
    # panel length relative to the longest, quantised so mirror images compare equal
    level = [round(ls[i] / out[i] / longest * 1.0e9) for i in range(len(out))]
    slack = max(2, int(math.ceil(0.01 * target)))
    for value in sorted(set(level), reverse=True):
        if left <= 0:
            break
        members = [i for i in range(len(out)) if level[i] == value]
        if len(members) > left + slack:
            members = sorted(members, key=_bit_reversed_fraction)[:left]
        for i in members:
            out[i] += 1
        left -= len(members)
    return out


# ----------------------------------------------------------------------------
# Surface-impedance laws of the snapshot's ``ibcs`` rows (twod/geometry.py:128-262, 261-312,
# 369-398, 603-680)
# ----------------------------------------------------------------------------
ETA0 = 376.730313668
TAPER_KINDS = ("constant", "linear", "cosine", "exp")
# A frequency this close (relative) to a table's first or last row samples that row.
TABLE_RANGE_RELATIVE_TOLERANCE = 1.0e-12
IMPEDANCE_CSV_HEADER = ["frequency_hz", "resistance_ohm", "reactance_ohm"]


def _passive(z: complex, context: str) -> complex:
    """A finite Leontovich impedance with Re(Zs) >= 0 up to 64 ulp (twod/geometry.py:513-526)."""
    z = complex(z)
    if not (math.isfinite(z.real) and math.isfinite(z.imag)):
        raise ValueError(f"{context} contains non-finite value {z!r}.")
    if z.real < -64.0 * np.finfo(float).eps * max(1.0, abs(z)):
        raise ValueError(f"{context} has negative resistance Re(Zs)={z.real:g} ohm. "
                         "Active/gain surface impedances are not supported.")
    return z


@dataclass
class ImpedanceLaw:
    """One IBC flag: an inline law along its segment (``constant``, ``linear``, ``cosine``, ``exp``
    between z_start at arc_s = 0 and z_end at arc_s = 1 in the drawing direction), a frequency
    table (``table``, linear interpolation in GHz) or a thin dielectric layer (``thin_dielectric``,
    a transmitting sheet that an opaque contour cannot carry)."""
    kind: str
    z_start: complex = 0j
    z_end: complex = 0j
    freqs_ghz: np.ndarray | None = None
    values: np.ndarray | None = None

    @property
    def tapered(self) -> bool:
        return self.kind in ("linear", "cosine", "exp")

    def _taper(self, arc_s: float) -> complex:
        s = float(max(0.0, min(1.0, arc_s)))
        if self.kind == "constant":
            return self.z_start
        if self.kind in ("linear", "cosine"):
            w = s if self.kind == "linear" else 0.5 * (1.0 - math.cos(math.pi * s))
            return (1.0 - w) * self.z_start + w * self.z_end
        z1 = self.z_start if abs(self.z_start) > EPS else complex(EPS, 0.0)
        z2 = self.z_end if abs(self.z_end) > EPS else complex(EPS, 0.0)
        return complex(np.exp((1.0 - s) * np.log(z1) + s * np.log(z2)))

    def _sample(self, frequency_ghz: float) -> complex:
        f = float(frequency_ghz)
        if not math.isfinite(f):
            raise ValueError("Material-table sample frequency must be finite.")
        lo, hi = float(self.freqs_ghz[0]), float(self.freqs_ghz[-1])
        if lo - TABLE_RANGE_RELATIVE_TOLERANCE * abs(lo) <= f < lo:
            f = lo
        elif hi < f <= hi + TABLE_RANGE_RELATIVE_TOLERANCE * abs(hi):
            f = hi
        elif f < lo or f > hi:
            raise ValueError(f"Material-table sample frequency {f:.17g} GHz is outside the characterized range "
                             f"[{lo:.17g}, {hi:.17g}] GHz.")
        if len(self.freqs_ghz) == 1:
            return complex(self.values[0])
        return complex(np.interp(f, self.freqs_ghz, self.values.real), np.interp(f, self.freqs_ghz, self.values.imag))

    def impedance(self, flag: int, frequency_ghz: float | None, arc_s: float | None = None) -> complex:
        """Zs of this law at ``frequency_ghz`` and the arc position ``arc_s`` (twod/geometry.py:369-390)."""
        if self.kind == "thin_dielectric":
            raise ValueError("A thin dielectric layer is a transmitting sheet, not an opaque IBC. "
                             "Assign it only to TYPE 1.")
        if self.kind == "table":
            return _passive(self._sample(frequency_ghz), f"IBC flag {flag} impedance sampled at {frequency_ghz:g} GHz")
        if self.tapered:
            s = 0.5 if arc_s is None else float(arc_s)
            return _passive(self._taper(s), f"IBC flag {flag} tapered impedance at s={s:g}")
        return _passive(self.z_start, f"IBC flag {flag} impedance")


def _material_flag(token, context: str = "IBC material definition flag") -> int:
    text = str(token).strip()
    if text.lower().startswith("mat."):
        text = text[4:]
    try:
        value = float(text)
    except (TypeError, ValueError):
        raise ValueError(f"{context} must be an integer; got {token!r}.") from None
    if not math.isfinite(value) or not value.is_integer():
        raise ValueError(f"{context} must be an integer; got {token!r}.")
    if value <= 0:
        raise ValueError(f"{context} must be a positive integer; got {token!r}.")
    return int(value)


def _material_float(token, context: str) -> float:
    try:
        value = float(token)
    except (TypeError, ValueError):
        raise ValueError(f"{context} must be a finite numeric value; got {token!r}.") from None
    if not math.isfinite(value):
        raise ValueError(f"{context} must be finite; got {token!r}.")
    return value


def material_base_dir(snapshot: dict, base_dir: str | None) -> str:
    """The one directory material tables are read from: the explicit one, else the folder of the
    snapshot's ``source_path``, else the working directory (twod/geometry.py:577-600)."""
    import os
    if base_dir is not None and str(base_dir).strip():
        return os.path.abspath(os.path.expanduser(str(base_dir)))
    source = str(snapshot.get("source_path", "") or "").strip()
    if source:
        return os.path.dirname(os.path.abspath(os.path.expanduser(source)))
    return os.path.abspath(os.getcwd())


def _table_file(folder: str, flag: int, filename) -> str:
    """A same-directory .csv sidecar name, resolved in ``folder`` only
    (ghost_backend/geometry/io.py:34-61, twod/geometry.py:566-575)."""
    import os
    raw = str(filename)
    name = raw.strip()
    context = f"Material flag {flag} CSV reference"
    if not name or not name.lower().endswith(".csv"):
        raise ValueError(f"{context} must name a .csv file; got {filename!r}.")
    if any(ch.isspace() for ch in raw):
        raise ValueError(f"{context} cannot contain whitespace because .geo material rows use unquoted fields; "
                         f"got {filename!r}. Rename the CSV and try again.")
    if name in (".", "..") or os.path.isabs(name) or os.path.basename(name) != name or "/" in name or "\\" in name:
        raise ValueError(f"{context} must be a filename in the same directory as the geometry file (no directory "
                         f"components); got {filename!r}.")
    path = os.path.join(os.path.abspath(folder), name)
    if not os.path.isfile(path):
        raise FileNotFoundError(f"Could not locate material file {name} in declared material directory "
                                f"{os.path.abspath(folder)}. Material tables are never searched in the process "
                                "working directory.")
    return path


def _read_csv_rows(path: str, header_names) -> list:
    """Numeric rows of a headered material CSV (UTF-8, optional BOM, # comments and blank lines
    skipped, frequencies in Hz positive and unique), sorted by frequency (twod/geometry.py:603-658)."""
    import csv
    names = list(header_names)
    if not str(path).lower().endswith(".csv"):
        raise ValueError(f"Material/IBC file must use the .csv extension: {path}")
    rows, header = [], False
    with open(path, "r", encoding="utf-8-sig", newline="") as handle:
        for lineno, raw in enumerate(handle, start=1):
            stripped = raw.strip()
            if not stripped or stripped.startswith("#"):
                continue
            try:
                parts = [part.strip() for part in next(csv.reader([raw], strict=True))]
            except csv.Error as exc:
                raise ValueError(f"{path}: line {lineno} contains invalid CSV: {exc}") from exc
            if not header:
                if parts != names:
                    raise ValueError(f"{path}: line {lineno} must have comma-separated header "
                                     f"{','.join(names)} (frequency in Hz); found {','.join(parts)}.")
                header = True
                continue
            if len(parts) != len(names):
                raise ValueError(f"{path}: line {lineno} must contain exactly {len(names)} "
                                 f"comma-separated columns; found {len(parts)}.")
            try:
                values = [float(part) for part in parts]
            except ValueError:
                raise ValueError(f"{path}: line {lineno} contains a non-numeric value.") from None
            if not all(math.isfinite(v) for v in values):
                raise ValueError(f"{path}: line {lineno} contains a non-finite value.")
            if values[0] <= 0:
                raise ValueError(f"{path}: line {lineno} has non-positive frequency {values[0]:g} Hz.")
            rows.append(values)
    if not header:
        raise ValueError(f"{path}: missing required header {','.join(names)}.")
    if not rows:
        raise ValueError(f"{path}: no material data rows found after the header.")
    rows.sort(key=lambda r: r[0])
    for a, b in zip(rows, rows[1:]):
        if a[0] == b[0]:
            raise ValueError(f"{path}: duplicate frequency {b[0]:g} Hz.")
    return rows


def load_impedance_csv(path: str) -> ImpedanceLaw:
    """``frequency_hz,resistance_ohm,reactance_ohm`` rows -> a table law (twod/geometry.py:661-680)."""
    rows = _read_csv_rows(path, IMPEDANCE_CSV_HEADER)
    values = np.array([complex(r[1], r[2]) for r in rows], dtype=np.complex128)
    for i, v in enumerate(values, start=1):
        _passive(v, f"Impedance CSV '{path}' data row {i}")
    return ImpedanceLaw("table", freqs_ghz=np.array([r[0] / 1.0e9 for r in rows]), values=values)


def impedance_laws(snapshot: dict, base_dir: str | None = None) -> dict:
    """flag -> ImpedanceLaw for every ``ibcs`` row of the snapshot (twod/geometry.py:261-312):
    ``[flag, kind, R_start, X_start, R_end, X_end]`` inline laws (a constant law ignores its end
    values), ``[flag, "file.csv"]`` frequency tables read from the material directory, and
    ``[flag, "thin_dielectric", ...]`` layers."""
    laws = {}
    folder = None
    for row in snapshot.get("ibcs", []) or []:
        if not row:
            continue
        flag = _material_flag(row[0])
        if flag in laws:
            raise ValueError(f"Duplicate IBC material flag {flag}.")
        if len(row) > 1 and str(row[1]).lower() == "thin_dielectric":
            laws[flag] = ImpedanceLaw("thin_dielectric")
            continue
        if len(row) == 2:
            if not str(row[1]).strip().lower().endswith(".csv"):
                raise ValueError(f"IBC flag {flag}: file-backed material definitions must use 'flag filename.csv'.")
            folder = folder or material_base_dir(snapshot, base_dir)
            laws[flag] = load_impedance_csv(_table_file(folder, flag, row[1]))
            continue
        tokens = [str(t).strip() for t in row[1:] if str(t).strip() != ""]
        if len(tokens) != 5:
            raise ValueError(f"IBC flag {flag}: inline impedance requires '<kind> R_start X_start R_end X_end' "
                             f"(got {len(tokens)} data tokens after the flag).")
        kind = tokens[0].lower()
        r0 = _material_float(tokens[1], f"IBC flag {flag} R_start")
        x0 = _material_float(tokens[2], f"IBC flag {flag} X_start")
        r1 = _material_float(tokens[3], f"IBC flag {flag} R_end")
        x1 = _material_float(tokens[4], f"IBC flag {flag} X_end")
        if kind not in TAPER_KINDS:
            raise ValueError(f"Unknown impedance taper kind '{kind}'. Expected one of {TAPER_KINDS}.")
        z0 = _passive(complex(r0, x0), "taper z_start")
        z1 = z0 if kind == "constant" else _passive(complex(r1, x1), "taper z_end")
        laws[flag] = ImpedanceLaw(kind, z0, z1)
    return laws


# ----------------------------------------------------------------------------
# Dielectric media of the snapshot's ``dielectrics`` rows (twod/geometry.py:149-160, 236-258,
# 314-367, 392-405, 530-564, 683-706, 2349-2394)
# ----------------------------------------------------------------------------
MEDIUM_CSV_HEADER = ["frequency_hz", "eps_real", "eps_imag", "mu_real", "mu_imag"]
MATERIAL_SINGULAR_TOL = EPS


def _finite(z: complex, context: str) -> complex:
    z = complex(z)
    if not (math.isfinite(z.real) and math.isfinite(z.imag)):
        raise ValueError(f"{context} contains non-finite value {z!r}.")
    return z


def passive_medium(eps: complex, mu: complex, context: str):
    """Finite, non-singular eps and mu with Im <= 0 up to 64 ulp (exp(+j omega t))."""
    eps = _finite(eps, f"{context} epsilon")
    mu = _finite(mu, f"{context} mu")
    if abs(eps) <= MATERIAL_SINGULAR_TOL:
        raise ValueError(f"{context} has unsupported singular/near-ENZ epsilon {eps!r} (|epsilon| <= "
                         f"{MATERIAL_SINGULAR_TOL:g}); it will not be replaced with free space.")
    if abs(mu) <= MATERIAL_SINGULAR_TOL:
        raise ValueError(f"{context} has unsupported singular/near-MNZ mu {mu!r} (|mu| <= "
                         f"{MATERIAL_SINGULAR_TOL:g}); it will not be replaced with free space.")
    for name, sym, v in (("epsilon", "epsilon", eps), ("mu", "mu", mu)):
        if v.imag > 64.0 * np.finfo(float).eps * max(1.0, abs(v)):
            raise ValueError(f"{context} {name} has gain-sign Im({sym})={v.imag:g}. For e^(+j*omega*t), passive "
                             f"media require Im({sym}) <= 0; active/gain media are not supported.")
    return eps, mu


def causal_index(eps: complex, mu: complex) -> complex:
    """Refractive index sqrt(eps mu) on the passive branch: Im n <= 0, and on the lossless branch
    the sign that gives Re(mu / n) >= 0."""
    n = complex(np.sqrt(complex(eps) * complex(mu)))
    tol = 64.0 * np.finfo(float).eps * max(1.0, abs(n))
    if n.imag > tol:
        n = -n
    elif abs(n.imag) <= tol and (complex(mu) / n).real < 0.0:
        n = -n
    return n


@dataclass
class Medium:
    """A dielectric flag: inline (eps, mu) or a frequency table of both."""
    eps: complex = 1.0 + 0.0j
    mu: complex = 1.0 + 0.0j
    freqs_ghz: np.ndarray | None = None
    eps_values: np.ndarray | None = None
    mu_values: np.ndarray | None = None

    def sample(self, flag: int, frequency_ghz: float):
        if self.freqs_ghz is None:
            return passive_medium(self.eps, self.mu, f"Dielectric flag {flag}")
        table = ImpedanceLaw("table", freqs_ghz=self.freqs_ghz, values=self.eps_values)
        eps = table._sample(frequency_ghz)
        mu = ImpedanceLaw("table", freqs_ghz=self.freqs_ghz, values=self.mu_values)._sample(frequency_ghz)
        return passive_medium(eps, mu, f"Dielectric flag {flag} sampled at {frequency_ghz:g} GHz")


def dielectric_media(snapshot: dict, base_dir: str | None = None) -> dict:
    """flag -> Medium for every ``dielectrics`` row: ``[flag, eps_real, eps_imag, mu_real, mu_imag]``
    or ``[flag, "file.csv"]`` (``frequency_hz,eps_real,eps_imag,mu_real,mu_imag``)."""
    media = {}
    folder = None
    for row in snapshot.get("dielectrics", []) or []:
        if not row:
            continue
        flag = _material_flag(row[0], "Dielectric material definition flag")
        if flag in media:
            raise ValueError(f"Duplicate dielectric material flag {flag}.")
        if len(row) == 2:
            if not str(row[1]).strip().lower().endswith(".csv"):
                raise ValueError(f"Dielectric flag {flag}: inline material requires exactly 'flag eps_real eps_imag "
                                 "mu_real mu_imag'; file-backed definitions require 'flag filename.csv'.")
            folder = folder or material_base_dir(snapshot, base_dir)
            path = _table_file(folder, flag, row[1])
            rows = _read_csv_rows(path, MEDIUM_CSV_HEADER)
            eps = np.array([complex(r[1], r[2]) for r in rows])
            mu = np.array([complex(r[3], r[4]) for r in rows])
            for i, (e, m) in enumerate(zip(eps, mu), start=1):
                passive_medium(e, m, f"Dielectric CSV '{path}' data row {i}")
            media[flag] = Medium(freqs_ghz=np.array([r[0] / 1.0e9 for r in rows]), eps_values=eps, mu_values=mu)
            continue
        if len(row) != 5 or any(str(t).strip() == "" for t in row):
            raise ValueError(f"Dielectric flag {flag}: inline material requires exactly 'flag eps_real eps_imag "
                             f"mu_real mu_imag' with no blank fields (got {len(row)} fields).")
        er = _material_float(row[1], f"Dielectric flag {flag} epsilon real part")
        ei = _material_float(row[2], f"Dielectric flag {flag} epsilon imaginary part")
        mr = _material_float(row[3], f"Dielectric flag {flag} mu real part")
        mi = _material_float(row[4], f"Dielectric flag {flag} mu imaginary part")
        eps = _finite(complex(er, ei), f"Dielectric flag {flag} epsilon")
        mu = _finite(complex(mr, mi), f"Dielectric flag {flag} mu")
        eps, mu = passive_medium(eps, mu, f"Dielectric flag {flag}")
        media[flag] = Medium(eps, mu)
    return media


def medium_of(media: dict, flag: int, frequency_ghz: float):
    if flag <= 0:
        return 1.0 + 0.0j, 1.0 + 0.0j
    if flag not in media:
        raise ValueError(f"Undefined dielectric flag {flag}.")
    return media[flag].sample(flag, frequency_ghz)


def material_wavelength(segments, media: dict, frequency_ghz: float) -> float:
    """Shortest wavelength the mesh of ``segments`` must resolve at one frequency: air's, divided
    by the largest |n| of the dielectric flags TYPE 3/4 (pos_mat) and TYPE 5 (both) segments
    reference (twod/geometry.py:2394-2452)."""
    flags = set()
    for seg in segments:
        props = list(seg.get("properties", []) or []) + [""] * 5
        seg_type = _parse_int(props[0]) if str(props[0]).strip() else int(seg.get("seg_type", 0))
        pos, neg = _parse_int(props[3]), _parse_int(props[4])
        if seg_type in (3, 4) and pos > 0:
            flags.add(pos)
        elif seg_type == 5:
            flags.update(f for f in (pos, neg) if f > 0)
    index = 1.0
    for flag in sorted(flags):
        value = abs(causal_index(*medium_of(media, flag, frequency_ghz)))
        if not math.isfinite(value) or value <= 0.0:
            raise ValueError(f"Dielectric flag {flag} produced invalid refractive-index magnitude {value!r} at "
                             f"{frequency_ghz:g} GHz.")
        index = max(index, value)
    return C0 / (float(frequency_ghz) * 1.0e9) / index


# ----------------------------------------------------------------------------
# Bound surface waves and junction grading (twod/geometry.py:737-905, 1468-1545)
# ----------------------------------------------------------------------------
SURFACE_WAVE_MAX_DENSITY_FACTOR = 4.0
SURFACE_WAVE_MAX_UNRESOLVED_FACTOR = 12.0
SURFACE_WAVE_TAPER_SAMPLES = 9
JUNCTION_GRADING_LEVELS = 4
JUNCTION_GRADING_RATIO = 0.5
JUNCTION_GRADING_CONTRAST = 4.0
JUNCTION_GRADING_DENSITY = 32.0


def bound_surface_wave_index(z: complex) -> float:
    """Phase index of the wave bound to a flat surface of impedance ``z`` in air: with
    kappa/k = -j Zs/eta0 (TE) or -j eta0/Zs (TM), a wave counts when Re kappa > 0 and
    |Im beta| <= Re beta for beta/k = sqrt(1 + (kappa/k)^2); the larger index of the two."""
    z = complex(z)
    if abs(z) <= EPS:
        return 0.0
    best = 0.0
    for ratio in (z / ETA0, ETA0 / z):
        if (-1j * ratio).real <= 0.0:
            continue
        beta = complex(np.sqrt(complex(1.0 - ratio * ratio)))
        if abs(beta.imag) > beta.real:
            continue
        best = max(best, abs(beta.real))
    return best


def surface_wave_mesh_wavelength(material_wavelength: float, law: ImpedanceLaw | None, flag: int,
                                 frequencies) -> float:
    """Mesh wavelength of an impedance segment: the shortest bound-wave wavelength over the mesh
    frequencies when the air mesh would leave it unresolved by at most 12x, never below a quarter
    of the air wavelength (twod/geometry.py:1517-1545)."""
    if law is None or not frequencies or law.kind == "thin_dielectric":
        return material_wavelength
    shortest = material_wavelength
    for f in frequencies:
        positions = np.linspace(0.0, 1.0, SURFACE_WAVE_TAPER_SAMPLES) if law.tapered else (0.5,)
        try:
            index = max(bound_surface_wave_index(law.impedance(flag, f, float(s))) for s in positions)
        except ValueError:
            index = 0.0
        if index > 0.0:
            bound = C0 / (float(f) * 1.0e9) / index
            if material_wavelength <= SURFACE_WAVE_MAX_UNRESOLVED_FACTOR * bound:
                shortest = min(shortest, bound)
    return max(shortest, material_wavelength / SURFACE_WAVE_MAX_DENSITY_FACTOR)


def junction_grading_levels(panel_length: float, wavelength: float) -> int:
    """Geometric levels toward a junction: 4, one more per doubling of the density beyond 32 panels
    per wavelength, the innermost panel kept above 1e-7 m."""
    levels = JUNCTION_GRADING_LEVELS
    if not panel_length > 0.0:
        return levels
    if math.isfinite(wavelength) and wavelength > 0.0:
        density = wavelength / panel_length
        if density > JUNCTION_GRADING_DENSITY:
            levels += int(math.ceil(math.log2(density / JUNCTION_GRADING_DENSITY) - 1.0e-9))
    return max(0, min(levels, int(math.floor(math.log2(max(panel_length / 1.0e-7, 1.0))))))


def impedance_jump_is_graded(first: complex, second: complex) -> bool:
    """Two evaluated laws (0 = PEC) differ by more than 3/4 of the larger magnitude."""
    larger = max(abs(first), abs(second))
    return larger > 0.0 and abs(first - second) > (1.0 - 1.0 / JUNCTION_GRADING_CONTRAST) * larger


class _Welder:
    """Vertex identity on the 1e-9 m snap: the snap key of a registered vertex, or the nearest
    registered vertex within the tolerance (twod/geometry.py:1994-2044)."""

    def __init__(self, tol: float = NODE_SNAP):
        self.tol = tol
        self.cell = 2.0 * max(tol, EPS)
        self.alias, self.grid = {}, {}

    def _raw(self, xy):
        return int(round(float(xy[0]) / self.tol)), int(round(float(xy[1]) / self.tol))

    def lookup(self, xy):
        known = self.alias.get(self._raw(xy))
        if known is not None:
            return known
        x, y = float(xy[0]), float(xy[1])
        cx, cy = math.floor(x / self.cell), math.floor(y / self.cell)
        best = None
        for i in (cx - 1, cx, cx + 1):
            for j in (cy - 1, cy, cy + 1):
                for px, py, key in self.grid.get((i, j), ()):
                    d = math.hypot(px - x, py - y)
                    if d <= self.tol and (best is None or d < best[0]):
                        best = (d, key)
        return None if best is None else best[1]

    def key(self, xy):
        raw = self._raw(xy)
        known = self.lookup(xy)
        if known is None:
            known = raw
            x, y = float(xy[0]), float(xy[1])
            self.grid.setdefault((math.floor(x / self.cell), math.floor(y / self.cell)), []).append((x, y, raw))
        self.alias.setdefault(raw, known)
        return known


def boundary_condition_junctions(segments, scale: float, laws: dict, frequencies):
    """Welded vertices where the evaluated law of the conductor segments ending there jumps at
    some mesh frequency (segment ends at arc_s 0 and 1 in the drawing direction; an unevaluable
    law next to PEC counts as a jump) (twod/geometry.py:825-866).  Returns (welder, keys)."""
    welder = _Welder()
    ends = {}
    freqs = list(frequencies) or [None]
    for seg in segments:
        props = list(seg.get("properties", []) or [])
        seg_type = _parse_int(props[0]) if props and str(props[0]).strip() else int(seg.get("seg_type", 2))
        pairs = list(seg.get("point_pairs", []) or [])
        if seg_type not in (2, 4) or not pairs:
            continue
        flag = _parse_int(props[2]) if len(props) > 2 else 0
        for pair, (kx, ky), arc in ((pairs[0], ("x1", "y1"), 0.0), (pairs[-1], ("x2", "y2"), 1.0)):
            key = welder.key((float(pair.get(kx, 0.0)) * scale, float(pair.get(ky, 0.0)) * scale))
            values = []
            for f in freqs:
                if flag <= 0:
                    values.append(0j)
                    continue
                law = laws.get(flag)
                if law is None or (f is None and law.kind == "table"):
                    values.append(None)
                    continue
                try:
                    z = law.impedance(flag, f, arc)
                except ValueError:
                    values.append(None)
                    continue
                values.append(0j if abs(z) <= EPS else z)
            ends.setdefault(key, []).append(values)
    junctions = set()
    for key, per_end in ends.items():
        for at_f in zip(*per_end):
            known = [v for v in at_f if v is not None]
            jump = len(known) < len(at_f) and any(v == 0.0 for v in known)
            if jump or any(impedance_jump_is_graded(a, b) for i, a in enumerate(known) for b in known[i + 1:]):
                junctions.add(key)
                break
    return welder, junctions


def grade_toward_junctions(pts, welder, junctions, wavelength: float):
    """Split the end panels of a primitive geometrically (ratio 1/2) toward junction vertices;
    a one-panel primitive with junctions at both ends is halved first (twod/geometry.py:869-891)."""
    if len(pts) < 2 or not junctions:
        return pts

    def at(xy):
        key = welder.lookup(xy)
        return key is not None and key in junctions

    at_start, at_end = at(pts[0]), at(pts[-1])
    if not (at_start or at_end):
        return pts
    if len(pts) == 2 and at_start and at_end:
        pts = [pts[0], 0.5 * (pts[0] + pts[1]), pts[1]]

    def fractions(first, second):
        levels = junction_grading_levels(float(np.linalg.norm(second - first)), wavelength)
        return [JUNCTION_GRADING_RATIO ** level for level in range(levels, 0, -1)]

    if at_start:
        pts = [pts[0]] + [pts[0] + (pts[1] - pts[0]) * f for f in fractions(pts[0], pts[1])] + pts[1:]
    if at_end:
        pts = pts[:-1] + [pts[-1] + (pts[-2] - pts[-1]) * f for f in reversed(fractions(pts[-1], pts[-2]))] + [pts[-1]]
    return pts


def build_mesh(snapshot: dict, units: str, frequency_ghz: float, max_panels: int = 20000, *,
               material_base_dir: str | None = None, law_ghz: float | None = None,
               mesh_frequencies=None) -> ContourMesh:
    """Panels and welded nodes of the conductor contours of ``snapshot`` meshed at
    ``frequency_ghz``, with each element's surface impedance evaluated at ``law_ghz`` (default:
    the mesh frequency).  ``mesh_frequencies`` (default: the mesh frequency) are the frequencies the
    mesh serves: junction grading and the bound-surface-wave mesh wavelength use them.

    Impedance segments take their law's value at each panel's arc-length centre along the
    segment in the drawing direction (tapers), at ``law_ghz`` (tables); junctions where the law
    jumps are graded (twod/geometry.py:1800-1981, 2600-2625).

    A snapshot that carries ``_2d_certification_refinement_factor`` (> 1) and
    ``_2d_certification_base_segment_n`` (each segment's base N) is the fine mesh of a
    certification pair: its chains get ``refined_chain_counts`` of the base counts
    (twod/geometry.py:1788-1850)."""
    scale = unit_scale(units)
    law_ghz = float(frequency_ghz if law_ghz is None else law_ghz)
    freqs = [float(frequency_ghz)] if mesh_frequencies is None else sorted({float(f) for f in mesh_frequencies})
    segments = list(snapshot.get("segments", []) or [])
    if not segments:
        raise ValueError("Geometry does not contain any segments.")
    refine = float(snapshot.get("_2d_certification_refinement_factor", 1.0) or 1.0)
    base_n = list(snapshot.get("_2d_certification_base_segment_n", []) or [])
    if not math.isfinite(refine) or refine < 1.0:
        raise ValueError("Internal 2-D certification refinement factor must be finite and >= 1.")
    if refine > 1.0 and len(base_n) != len(segments):
        raise ValueError("Internal 2-D certification refinement metadata does not match the geometry "
                         "segment count.")
    laws = impedance_laws(snapshot, material_base_dir)
    media = dielectric_media(snapshot, material_base_dir)
    wavelength0 = min(material_wavelength(segments, media, f) for f in freqs)
    welder, junctions = boundary_condition_junctions(segments, scale, laws, freqs)
    panels0, panels1, zs, keys_e = [], [], [], []
    types_e, flags_e, names_e = [], [], []
    kinds, dielectric_flags = set(), set()
    for s, seg in enumerate(segments):
        seg_type, n_prop, ibc = _parse_props(seg)
        name = str(seg.get("name", "segment"))
        if seg_type not in (1, 2, 3, 4, 5):
            raise ValueError(f"Segment '{name}': TYPE {seg_type} boundaries are not supported by the GPU path: "
                             "sheets (TYPE 1), conductors (TYPE 2), dielectric interfaces with air (TYPE 3), "
                             "coated conductors (TYPE 4) and dielectric/dielectric interfaces (TYPE 5) are.")
        if seg_type == 1 and ibc <= 0:
            raise ValueError(f"TYPE 1 panel '{name}' requires IBC > 0 in coupled dielectric mode.")
        kinds.add(seg_type)
        # an explicit count is held to the media on the segment's own sides
        local_wavelength = min(material_wavelength([seg], media, f) for f in freqs)
        props5 = list(seg.get("properties", []) or []) + [0] * 5
        pos_mat, neg_mat = _parse_int(props5[3]), _parse_int(props5[4])
        if seg_type in (3, 4) and pos_mat <= 0:
            raise ValueError(f"TYPE {seg_type} panel '{name}' requires pos_mat > 0.")
        if seg_type == 5 and (pos_mat <= 0 or neg_mat <= 0):
            raise ValueError(f"TYPE 5 panel '{name}' requires pos_mat > 0 and neg_mat > 0.")
        if seg_type in (3, 5) and ibc > 0:
            raise ValueError(f"Segment '{name}': an impedance sheet on a dielectric interface is not supported by "
                             "the GPU path.")
        if seg_type == 3:
            dielectric_flags.add(pos_mat)
        # (minus region, plus region) of the segment's panels: 0 air, -1 conductor
        region_key = {1: (0, 0), 2: (0, -1), 3: (0, pos_mat), 4: (-1, pos_mat), 5: (neg_mat, pos_mat)}[seg_type]
        law = None
        if ibc > 0:
            if ibc not in laws:
                raise ValueError(f"Undefined IBC flag {ibc}.")
            law = laws[ibc]
            if law.kind == "thin_dielectric":
                if seg_type != 1:
                    raise ValueError("Thin dielectric layers can be assigned only to a free TYPE 1 sheet.")
                raise ValueError(f"Segment '{name}': transmitting thin dielectric layers on TYPE 1 sheets are not "
                                 "supported by the GPU path; resistive and reactive sheets (impedance laws) are.")
        # bound surface waves size opaque conductor segments only (twod/geometry.py:765, 1540-1545)
        wavelength = wavelength0 if seg_type == 1 else surface_wave_mesh_wavelength(wavelength0, law, ibc, freqs)
        pairs = list(seg.get("point_pairs", []) or [])
        if seg_type in (2, 3):
            # TYPE 2 and 3: the user draws the normal into air; reverse the chain so that the stored
            # normal points to the plus side (TYPE 4 and 5 are drawn toward their plus side).
            pairs = [{"x1": p.get("x2", 0.0), "y1": p.get("y2", 0.0), "x2": p.get("x1", 0.0), "y2": p.get("y1", 0.0)}
                     for p in reversed(pairs)]
        ends = []
        for pair in pairs:
            a = np.array([float(pair["x1"]) * scale, float(pair["y1"]) * scale])
            b = np.array([float(pair["x2"]) * scale, float(pair["y2"]) * scale])
            length = float(np.linalg.norm(b - a))
            if EPS < length <= NODE_SNAP:
                raise ValueError(f"Segment '{name}' has a primitive {length:.3g} m long, not longer than the "
                                 "mesh node tolerance.")
            ends.append((a, b, length))
        if refine > 1.0:
            seg_base_n = _parse_int(base_n[s])
            base_counts = [_panel_count(seg_base_n, length, wavelength) for _, _, length in ends]
            fine_total = max(sum(base_counts) + 1, int(math.ceil(sum(base_counts) * refine)))
            if len(panels0) + fine_total > max_panels:
                raise ValueError("Discretization exceeds the configured panel limit.")
            counts = refined_chain_counts(base_counts, [e[2] for e in ends], refine) if ends else []
        else:
            counts = [_panel_count(n_prop, length, wavelength) for _, _, length in ends]
        seg_start = len(panels0)
        for (a, b, length), count in zip(ends, counts):
            if n_prop > 0:
                minimum = max(1, int(math.ceil(length * MIN_EXPLICIT_PANELS_PER_WAVELENGTH / local_wavelength)))
                if count < minimum:
                    raise ValueError(f"Segment '{name}' explicit N={n_prop} under-resolves a {length:.6g} m "
                                     f"primitive at wavelength {local_wavelength:.6g} m: at least N={minimum} is "
                                     "required.")
            if len(panels0) + count > max_panels:
                raise ValueError("Discretization exceeds the configured panel limit.")
            if count > 1 and length / count <= NODE_SNAP:
                raise ValueError(f"Segment '{name}' N={n_prop} would split a {length:.3g} m primitive into "
                                 "panels no longer than the mesh node tolerance. Reduce N.")
            pts = [a + (b - a) * (i / count) for i in range(count + 1)]
            pts[0], pts[-1] = a.copy(), b.copy()
            if junctions:
                pts = grade_toward_junctions(pts, welder, junctions, wavelength)
                if len(panels0) + len(pts) - 1 > max_panels:
                    raise ValueError("Discretization exceeds the configured panel limit.")
            for i in range(len(pts) - 1):
                if np.linalg.norm(pts[i + 1] - pts[i]) <= EPS:
                    continue
                panels0.append(pts[i])
                panels1.append(pts[i + 1])
        # arc-length centre of each panel along the segment, in the drawing direction
        seg_len = [float(np.linalg.norm(panels1[i] - panels0[i])) for i in range(seg_start, len(panels0))]
        total, cum = sum(seg_len), 0.0
        for length in seg_len:
            arc = (cum + 0.5 * length) / total if total > EPS else 0.5
            cum += length
            zs.append(0j if law is None else law.impedance(ibc, law_ghz, 1.0 - arc if seg_type in (2, 3) else arc))
            keys_e.append(region_key)
            types_e.append(seg_type)
            flags_e.append(ibc)
            names_e.append(name.strip() or "__type1_sheet__")
    if not panels0:
        raise ValueError("Geometry does not contain any valid discretized panels.")
    p0 = np.array(panels0)
    p1 = np.array(panels1)
    if 1 in kinds:
        return SheetLayout(p0, p1, np.array(zs, complex), np.array(types_e, np.int64), np.array(flags_e, np.int64),
                           names_e)
    # node welding: first appearance, snap key or within the tolerance
    keys: dict = {}
    nodes: list = []
    ids = np.zeros((len(p0), 2), np.int64)
    for e in range(len(p0)):
        for c, xy in enumerate((p0[e], p1[e])):
            key = (int(round(xy[0] / NODE_SNAP)), int(round(xy[1] / NODE_SNAP)))
            found = keys.get(key)
            if found is None:
                for dx in (-2, -1, 0, 1, 2):
                    for dy in (-2, -1, 0, 1, 2):
                        idx = keys.get((key[0] + dx, key[1] + dy))
                        if idx is not None and np.linalg.norm(nodes[idx] - xy) <= NODE_SNAP:
                            found = idx
                            break
                    if found is not None:
                        break
            if found is None:
                found = len(nodes)
                nodes.append(xy.copy())
            keys.setdefault(key, found)
            ids[e, c] = found
    n = len(p0)
    first = ids[:, 0]
    if len(nodes) != n or sorted(first.tolist()) != list(range(n)):
        raise ValueError("The GPU path needs closed conductor contours drawn as chained primitives "
                         "(every node shared by exactly two panels).")
    if sorted(ids[:, 1].tolist()) != list(range(n)):
        raise ValueError("Every conductor contour must be closed.")
    # Element e starts at node e here; the reference numbers nodes by first appearance, which
    # differs where a contour is drawn as several segments.  ref_node keeps that numbering.
    ours = np.empty(n, np.int64)
    ours[first] = np.arange(n)
    nxt = ours[ids[:, 1]].astype(np.int32)
    prev = np.empty(n, np.int32)
    prev[nxt] = np.arange(n, dtype=np.int32)
    starts = [int(e) for e in range(n) if prev[e] != e - 1]
    impedance = np.array(zs, complex)
    has_ibc = bool(np.any(np.abs(impedance) > EPS))
    medium = regions = region_media = None
    if kinds == {3} and len(dielectric_flags) == 1:
        flag = dielectric_flags.pop()
        medium = (flag,) + tuple(medium_of(media, flag, law_ghz))
    elif kinds != {2}:
        # several regions (twod/solver.py:1850-1865): coated conductors, conductors beside
        # dielectrics, several media
        regions = np.array(keys_e, np.int64)
        used = sorted({int(r) for r in regions.ravel() if r > 0})
        if not np.any(regions == 0):
            raise ValueError("The exterior region must be air: a layout needs TYPE 2 or TYPE 3 boundaries.")
        region_media = {0: (1.0 + 0.0j, 1.0 + 0.0j)}
        region_media.update({r: tuple(medium_of(media, r, law_ghz)) for r in used})
    mesh = ContourMesh(np.array(nodes)[first], p0, p1, nxt, prev, starts,
                       impedance if has_ibc else None,
                       None if np.array_equal(first, np.arange(n)) else first.copy(), medium, regions, region_media)
    if regions is not None:
        _check_region_contours(mesh)
        if has_ibc:
            # one law per closed contour (PEC contours and impedance contours of several laws may share
            # a region); zones and tapers along a contour are the reference's element-weighted Robin
            # rows, not built here
            for cycle in _contours(mesh):
                if len({complex(impedance[e]) for e in cycle}) > 1:
                    raise ValueError("Surface impedances in multi-region layouts are supported by the GPU path as "
                                     "one constant law per closed contour; impedance zones and tapers along a "
                                     "contour are not.")
    _check_orientation(mesh)
    return mesh


def _contours(mesh: ContourMesh):
    seen = np.zeros(mesh.n, bool)
    for s in range(mesh.n):
        if seen[s]:
            continue
        cycle, e = [], s
        while not seen[e]:
            seen[e] = True
            cycle.append(e)
            e = int(mesh.next[e])
        yield cycle


def _check_region_contours(mesh: ContourMesh) -> None:
    """In a multi-region layout every closed contour carries one boundary kind: junctions between
    boundary kinds along a contour (the reference's junction conditions) are not built here."""
    for cycle in _contours(mesh):
        keys = {tuple(mesh.regions[e]) for e in cycle}
        if len(keys) != 1:
            raise ValueError("Junctions between boundary kinds along one contour are not supported by the GPU "
                             "path in multi-region layouts: each closed contour must carry one boundary kind.")


def _check_orientation(mesh: ContourMesh) -> None:
    """Each closed contour's stored normals must point into the body it bounds: a conductor or
    dielectric in air (TYPE 2, 3; drawn with the normal into air, reversed when stored) runs
    counterclockwise as stored; the inner boundary of a coating or of a layered body (TYPE 4, 5;
    drawn and stored with the normal out of the inner region) runs clockwise."""
    for cycle in _contours(mesh):
        area = 0.0
        for e in cycle:
            a, b = mesh.p0[e], mesh.p1[e]
            area += a[0] * b[1] - b[0] * a[1]
        inner = mesh.regions is not None and mesh.regions[cycle[0]][0] != 0
        if inner and area >= 0:
            raise ValueError("Geometry orientation check failed: a TYPE 4 or 5 contour must be drawn with its "
                             "normal from the inner region into the pos_mat region (clockwise around the core).")
        if not inner and area <= 0:
            raise ValueError("Geometry orientation check failed: a TYPE 2 or 3 contour must be drawn with its "
                             "normal into the air (clockwise around the body).")


# ----------------------------------------------------------------------------
# Sheets (TYPE 1 open contours), alone or beside PEC contours
# ----------------------------------------------------------------------------
SHEET_MIXED_MESSAGE = (
    "Mixed TYPE 1 sheet + (IBC-coated body / dielectric body / layered coating) geometries are not currently "
    "supported. Supported options: all-sheet (any number of TYPE 1 sheets, each with its own Z_s), or sheet + "
    "pure-PEC TYPE 2 body (mixed sheet+PEC is handled by _solve_mixed_sheet_pec).  For a tapered resistive "
    "treatment on a coated body, use TYPE 2 with a tapered IBC row -- that's the physically correct model for "
    "'resistance transitioning from air to a conducting body'.")
SHEET_ATTACHED_MESSAGE = ("TE sheet attached to a PEC body requires a coupled junction condition; this geometry is "
                          "not supported. Model the fin as a closed finite-thickness PEC contour.")
SHEET_BRANCH_MESSAGE = "TE sheet branching junctions require an unsupported junction condition."


class SheetMesh:
    """Linear mesh of one polarization of a sheet layout: element e runs from node elements[e, 0] to
    node elements[e, 1], nodes numbered by first appearance and shared only between panels of one
    interface signature (twod/geometry.py:2189-2290), so a node may have one element (a strip end),
    two, or more.  ``endpoints`` are the TE strip ends whose rows are pinned."""

    degree = 1

    def __init__(self, nodes, p0, p1, elements, impedance, sheet, endpoints):
        self.nodes = np.asarray(nodes, float)
        self.p0 = np.ascontiguousarray(p0, float)
        self.p1 = np.ascontiguousarray(p1, float)
        self.elements = np.asarray(elements, np.int64)
        self.impedance = np.asarray(impedance, complex)
        self.sheet = np.asarray(sheet, bool)
        self.endpoints = np.asarray(endpoints, np.int64)

    @property
    def n(self) -> int:
        return int(len(self.nodes))

    @property
    def ne(self) -> int:
        return int(len(self.p0))

    @property
    def lengths(self) -> np.ndarray:
        return np.linalg.norm(self.p1 - self.p0, axis=1)

    @property
    def normals(self) -> np.ndarray:
        t = (self.p1 - self.p0) / self.lengths[:, None]
        return np.stack((-t[:, 1], t[:, 0]), axis=1)

    def far_order(self, k: float, hypersingular: bool = True) -> int:
        """Far rule: the W-calibrated 8 / 16 points of the hypersingular traversal (twod/operators.py:
        1502-1527), 8 for the single layer alone (its graded rule never exceeds the configured 8)."""
        if not hypersingular:
            return 8
        lengths = self.lengths
        lmax, lmin = float(lengths.max()), float(lengths.min())
        return 8 if 1e-4 <= abs(k) * lmax <= 3.0 and lmin >= 1e-3 * lmax else 16

    def owners(self):
        """(own_ptr, own_el, own_loc): the elements of each node in CSR form, ascending element order."""
        pairs = sorted((int(self.elements[e, c]), e, c) for e in range(self.ne) for c in (0, 1))
        ptr = np.zeros(self.n + 1, np.int32)
        for nid, _, _ in pairs:
            ptr[nid + 1] += 1
        return (np.cumsum(ptr).astype(np.int32), np.array([e for _, e, _ in pairs], np.int32),
                np.array([c for _, _, c in pairs], np.int32))

    def next_ids(self) -> np.ndarray:
        """Id of each element's end in the near test's convention: the first element starting at that
        node, else an id >= ne of the node alone."""
        starts = {}
        for e in range(self.ne):
            starts.setdefault(int(self.elements[e, 0]), e)
        out = np.empty(self.ne, np.int32)
        for e in range(self.ne):
            end = int(self.elements[e, 1])
            out[e] = starts.get(end, self.ne + end)
        return out


def _closed_conductor_loops(p0, p1, body) -> set:
    """Panels of closed one-sided conductor loops (their impedance flags do not split nodes,
    twod/geometry.py:2143-2186)."""
    welder = _Welder()
    ends = {i: (welder.key(p0[i]), welder.key(p1[i])) for i in np.flatnonzero(body)}
    incidence: dict = {}
    for i, keys in ends.items():
        for key in keys:
            incidence.setdefault(key, []).append(i)
    closed, remaining = set(), set(ends)
    while remaining:
        component, stack = set(), [remaining.pop()]
        while stack:
            i = stack.pop()
            if i in component:
                continue
            component.add(i)
            for key in ends[i]:
                stack.extend(j for j in incidence[key] if j not in component)
        remaining -= component
        if all(len(incidence[key]) == 2 for i in component for key in ends[i]):
            closed |= component
    return closed


def _orient_sheets(p0, p1, sheet):
    """TE: walk each non-branching sheet chain from a free end (its lowest-index panel's first free
    end, else that panel's start) and flip panels to follow the walk (twod/geometry.py:2292-2343)."""
    p0, p1 = p0.copy(), p1.copy()
    welder = _Welder()
    ends, incidence = {}, {}
    for i in np.flatnonzero(sheet):
        ends[i] = (welder.key(p0[i]), welder.key(p1[i]))
        for key in ends[i]:
            incidence.setdefault(key, []).append(int(i))
    if any(len(v) > 2 for v in incidence.values()):
        raise ValueError(SHEET_BRANCH_MESSAGE)
    remaining = {i for v in incidence.values() for i in v}
    flipped = np.zeros(len(p0), bool)
    while remaining:
        first = min(remaining)
        component, stack = set(), [first]
        while stack:
            i = stack.pop()
            if i in component:
                continue
            component.add(i)
            for key in ends[i]:
                stack.extend(j for j in incidence[key] if j not in component)
        start = ends[first][0]
        found = False
        for i in sorted(component):
            for key in ends[i]:
                if len(incidence[key]) == 1:
                    start, found = key, True
                    break
            if found:
                break
        while True:
            choices = [i for i in incidence[start] if i in remaining]
            if not choices:
                break
            i = min(choices)
            head, tail = ends[i]
            if head != start:
                p0[i], p1[i] = p1[i].copy(), p0[i].copy()
                flipped[i] = True
                head, tail = tail, head
            remaining.discard(i)
            start = tail
    return p0, p1, flipped


def _sheet_endpoints(nodes, p0, p1, elements, sheet, keys) -> np.ndarray:
    """Strip ends: nodes that one sheet panel end lands on (counted by welded vertex).  Raises for a
    sheet that touches a body, which needs a coupled junction condition
    (twod/solver.py:1683-1736)."""
    count, sheet_nodes, body_keys = {}, set(), set()
    for e in range(len(p0)):
        if not sheet[e]:
            body_keys.update(keys[int(n)] for n in elements[e])
            continue
        for n in elements[e]:
            sheet_nodes.add(int(n))
            count[keys[int(n)]] = count.get(keys[int(n)], 0) + 1
    attached = bool(body_keys.intersection(count))
    if body_keys and not attached:
        starts, dirs = p0[~sheet], p1[~sheet] - p0[~sheet]
        len2 = np.sum(dirs * dirs, axis=1)
        for n in sheet_nodes:
            if count[keys[n]] != 1:
                continue
            rel = nodes[n] - starts
            t = np.clip(np.sum(rel * dirs, axis=1) / len2, 0.0, 1.0)
            if np.any(np.linalg.norm(rel - t[:, None] * dirs, axis=1) <= 1e-9):
                attached = True
                break
    if attached:
        raise ValueError(SHEET_ATTACHED_MESSAGE)
    return np.array(sorted(n for n in sheet_nodes if count[keys[n]] == 1), np.int64)


class SheetLayout:
    """Panels of a geometry with TYPE 1 sheets: all sheets, or sheets beside pure-PEC TYPE 2 contours
    (twod/solver.py:1557-1633, 1739-1777).  The two polarizations mesh it differently: TE orients every
    sheet chain and lets sheets share nodes whatever their impedance flag or name; TM shares a node
    only between panels of one flag and one sheet name (twod/geometry.py:2189-2290).  ``mesh(pol)``
    builds that polarization's SheetMesh."""

    degree = 1
    medium = None
    regions = None

    def __init__(self, p0, p1, impedance, types, flags, names):
        self.p0, self.p1 = np.asarray(p0, float), np.asarray(p1, float)
        self.impedance = np.asarray(impedance, complex)
        self.types, self.flags, self.names = np.asarray(types), np.asarray(flags), list(names)
        self.sheet = self.types == 1
        body = ~self.sheet
        if np.any(body & ((self.types != 2) | (np.abs(self.impedance) > EPS))):
            raise ValueError(SHEET_MIXED_MESSAGE)
        for e in np.flatnonzero(self.sheet):
            if abs(self.impedance[e]) <= EPS:
                raise ValueError(f"TYPE 1 panel '{self.names[e]}' has zero impedance; provide non-zero IBC for "
                                 "sheet mode.")
        self.closed = _closed_conductor_loops(self.p0, self.p1, body)
        self._meshes = {}

    @property
    def n(self) -> int:
        """Panel count."""
        return int(len(self.p0))

    @property
    def lengths(self) -> np.ndarray:
        return np.linalg.norm(self.p1 - self.p0, axis=1)

    @property
    def all_sheet(self) -> bool:
        return bool(np.all(self.sheet))

    def mesh(self, pol: str) -> SheetMesh:
        pol = pol.upper()
        if pol not in self._meshes:
            self._meshes[pol] = self._build(pol)
        return self._meshes[pol]

    def _build(self, pol: str) -> SheetMesh:
        te = pol == "TE"
        p0, p1 = (_orient_sheets(self.p0, self.p1, self.sheet)[:2] if te else (self.p0.copy(), self.p1.copy()))
        welder = _Welder()
        index, nodes, keys = {}, [], []
        elements = np.zeros((len(p0), 2), np.int64)
        for e in range(len(p0)):
            if self.sheet[e]:
                sig = (1, 0, "") if te else (1, int(self.flags[e]), self.names[e])
            else:
                sig = (2, 0 if e in self.closed else int(self.flags[e]), "")
            for c, xy in enumerate((p0[e], p1[e])):
                key = welder.key(xy)
                found = index.get((key, sig))
                if found is None:
                    found = len(nodes)
                    index[(key, sig)] = found
                    nodes.append(np.array(xy, float))
                    keys.append(key)
                elements[e, c] = found
        nodes = np.array(nodes)
        endpoints = (_sheet_endpoints(nodes, p0, p1, elements, self.sheet, keys) if te else np.zeros(0, np.int64))
        return SheetMesh(nodes, p0, p1, elements, self.impedance, self.sheet, endpoints)
