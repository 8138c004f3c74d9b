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

    @property
    def n(self) -> int:
        return len(self.p0)

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
        """W-calibrated far rule of the paired traversal (twod/operators.py:1509-1527)."""
        lengths = self.lengths
        lmax, lmin = float(lengths.max()), float(lengths.min())
        if 1e-4 <= abs(k) * lmax <= 3.0 and lmin >= 1e-3 * lmax:
            return 8
        return 16


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


def impedance_laws(snapshot: dict) -> dict:
    """Inline constant surface impedances of the snapshot's ``ibcs`` rows,
    ``[flag, "constant", R, X, R_end, X_end]`` (twod/geometry.py:261-312: inline rows are
    ``<kind> R_start X_start R_end X_end``).  Tapered, file-backed and thin-layer laws are not
    handled by this path and are refused when a segment uses them."""
    laws = {}
    for row in snapshot.get("ibcs", []) or []:
        if not row:
            continue
        try:
            flag = int(round(float(row[0])))
        except (TypeError, ValueError):
            raise ValueError(f"Invalid IBC material definition flag {row[0]!r}.")
        if flag in laws:
            raise ValueError(f"Duplicate IBC material flag {flag}.")
        tokens = [str(t).strip() for t in row[1:] if str(t).strip() != ""]
        if len(tokens) != 5 or tokens[0].lower() != "constant":
            laws[flag] = None            # a law this path does not evaluate
            continue
        try:
            z = complex(float(tokens[1]), float(tokens[2]))
        except ValueError:
            raise ValueError(f"IBC flag {flag}: impedance values must be numeric.")
        if not (math.isfinite(z.real) and math.isfinite(z.imag)):
            raise ValueError(f"IBC flag {flag}: impedance must be finite.")
        if z.real < -1e-12 * max(1.0, abs(z)):
            raise ValueError(f"IBC flag {flag} has negative resistance Re(Zs)={z.real:g} ohm. Active/gain "
                             "surface impedances are not supported.")
        laws[flag] = 0j if abs(z) <= EPS else z
    return laws


def build_mesh(snapshot: dict, units: str, frequency_ghz: float, max_panels: int = 20000) -> ContourMesh:
    """Panels and welded nodes of the PEC contours of ``snapshot`` at ``frequency_ghz``.

    A snapshot that carries ``_2d_certification_refinement_factor`` (> 1) and
    ``_2d_certification_base_segment_n`` (each segment's base N) is the fine mesh of a
    certification pair: its chains get ``refined_chain_counts`` of the base counts
    (twod/geometry.py:1788-1850)."""
    scale = unit_scale(units)
    wavelength = C0 / (float(frequency_ghz) * 1e9)
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
    laws = impedance_laws(snapshot)
    panels0, panels1, zs = [], [], []
    for s, seg in enumerate(segments):
        seg_type, n_prop, ibc = _parse_props(seg)
        name = str(seg.get("name", "segment"))
        if seg_type != 2:
            raise ValueError(f"Segment '{name}': only conductor (TYPE 2) contours are supported by the GPU path.")
        z_seg = 0j
        if ibc > 0:
            if ibc not in laws:
                raise ValueError(f"Segment '{name}' uses IBC flag {ibc}, which the snapshot does not define.")
            if laws[ibc] is None:
                raise ValueError(f"Segment '{name}': only constant inline surface impedances are supported by "
                                 "the GPU path (no tapers, tables or thin layers).")
            z_seg = laws[ibc]
            if z_seg != 0 and n_prop <= 0:
                raise ValueError(f"Segment '{name}': impedance segments need an explicit N > 0 here (the automatic "
                                 "density follows the bound surface wave).")
        pairs = list(seg.get("point_pairs", []) or [])
        # TYPE 2: the user draws the normal into air; reverse the chain.
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
        for (a, b, length), count in zip(ends, counts):
            if n_prop > 0:
                minimum = max(1, int(math.ceil(length * MIN_EXPLICIT_PANELS_PER_WAVELENGTH / wavelength)))
                if count < minimum:
                    raise ValueError(f"Segment '{name}' explicit N={n_prop} under-resolves a {length:.6g} m "
                                     f"primitive at wavelength {wavelength:.6g} m: at least N={minimum} is "
                                     "required.")
            if len(panels0) + count > max_panels:
                raise ValueError("Discretization exceeds the configured panel limit.")
            if count > 1 and length / count <= NODE_SNAP:
                raise ValueError(f"Segment '{name}' N={n_prop} would split a {length:.3g} m primitive into "
                                 "panels no longer than the mesh node tolerance. Reduce N.")
            pts = [a + (b - a) * (i / count) for i in range(count + 1)]
            pts[0], pts[-1] = a.copy(), b.copy()
            for i in range(count):
                if np.linalg.norm(pts[i + 1] - pts[i]) <= EPS:
                    continue
                panels0.append(pts[i])
                panels1.append(pts[i + 1])
                zs.append(z_seg)
    if not panels0:
        raise ValueError("Geometry does not contain any valid discretized panels.")
    p0 = np.array(panels0)
    p1 = np.array(panels1)
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
    if len(nodes) != n or not np.array_equal(ids[:, 0], np.arange(n)):
        raise ValueError("The GPU path needs closed PEC contours drawn as chained primitives "
                         "(every node shared by exactly two panels).")
    nxt = ids[:, 1].astype(np.int32)
    if sorted(nxt.tolist()) != list(range(n)):
        raise ValueError("Every PEC contour must be closed.")
    prev = np.empty(n, np.int32)
    prev[nxt] = np.arange(n, dtype=np.int32)
    starts = [int(e) for e in range(n) if prev[e] != e - 1]
    impedance = np.array(zs, complex)
    mesh = ContourMesh(np.array(nodes), p0, p1, nxt, prev, starts,
                       impedance if np.any(impedance != 0) else None)
    _check_orientation(mesh)
    if mesh.impedance is not None:
        _check_uniform_law(mesh)
    return mesh


def _check_uniform_law(mesh: ContourMesh) -> None:
    """Each closed contour carries one law (PEC or one constant impedance): the flux rows
    Z_TE + alpha Z_TM hold exactly then; zones along a contour need element-weighted terms."""
    seen = np.zeros(mesh.n, bool)
    for s in range(mesh.n):
        if seen[s]:
            continue
        e, law = s, mesh.impedance[s]
        while not seen[e]:
            seen[e] = True
            if mesh.impedance[e] != law:
                raise ValueError("Impedance zones along one closed contour are not supported by the GPU path: "
                                 "each contour must be PEC or carry one constant surface impedance.")
            e = int(mesh.next[e])


def _check_orientation(mesh: ContourMesh) -> None:
    """Each closed contour must run counterclockwise internally (normals into the conductor)."""
    seen = np.zeros(mesh.n, bool)
    for s in range(mesh.n):
        if seen[s]:
            continue
        area, e = 0.0, s
        while not seen[e]:
            seen[e] = True
            a, b = mesh.p0[e], mesh.p1[e]
            area += a[0] * b[1] - b[0] * a[1]
            e = int(mesh.next[e])
        if area <= 0:
            raise ValueError("Geometry orientation check failed: a TYPE 2 contour must be drawn with its "
                             "normal into the air (clockwise around the body).")
