"""Host-side, k-independent set-up of the port-solution path.

Discretisation (reference bspline.py:988-1093), clamped B-spline bases per
wire with free/junction end handling and KCL rows (bspline.py:1216-1414), the
port drive/readout columns (bspline.py:2806-2870, 3784-3867) and the tables
the fill kernel reads (segment quadrature data, band polynomials, basis tiles,
same-edge table layout).  Everything here depends only on the geometry, so a
solver builds it once and reuses it for every frequency, as the reference
caches geometry and basis (bspline.py:192-195, 1004-1014, 1260-1264).

The basis polynomials are computed with a Cox-de Boor recursion carried out on
polynomials in the segment-local coordinate (the reference samples
scipy.interpolate.BSpline.design_matrix at d+1 points and inverts a
Vandermonde matrix instead, bspline.py:1316-1360); both give the same
coefficients to rounding.
"""

from dataclasses import dataclass, field

import numpy as np

from ._lib import TB, TSMAX

EPS0 = 8.8541878188e-12  # reference bspline.py:480
MU0 = 1.25663706127e-6   # reference bspline.py:481


# ---------------------------------------------------------------------------
# B-spline polynomials by Cox-de Boor on polynomial coefficients
# ---------------------------------------------------------------------------
def _poly_mul_linear(c, a0, a1):
    """(a0 + a1 u) * sum c_k u^k (ascending coefficient arrays, last axis)."""
    out = np.zeros(c.shape[:-1] + (c.shape[-1] + 1,))
    out[..., :-1] += a0[..., None] * c
    out[..., 1:] += a1[..., None] * c
    return out


def span_polynomials(knots, d, span, x0):
    """Coefficients (in u = x - x0) of the d+1 B-splines nonzero on knot span
    ``span`` (t[span] <= x < t[span+1]); returns (d+1 bases, d+1 powers),
    row b = basis span-d+b."""
    t = knots
    # degree 0: only B_span is 1 on this span
    basis = {span: np.array([1.0])}
    for k in range(1, d + 1):
        nxt = {}
        for j in range(span - k, span + 1):
            c = np.zeros(k + 1)
            if j in basis:  # (x - t_j)/(t_{j+k} - t_j) B_{j,k-1}
                den = t[j + k] - t[j]
                if den > 0:
                    c += _poly_mul_linear(basis[j], np.float64((x0 - t[j]) / den),
                                          np.float64(1.0 / den))
            if j + 1 in basis:  # (t_{j+k+1} - x)/(t_{j+k+1} - t_{j+1}) B_{j+1,k-1}
                den = t[j + k + 1] - t[j + 1]
                if den > 0:
                    c += _poly_mul_linear(basis[j + 1], np.float64((t[j + k + 1] - x0) / den),
                                          np.float64(-1.0 / den))
            nxt[j] = c
        basis = nxt
    return np.stack([basis[j] for j in range(span - d, span + 1)])


def bspline_values(knots, d, n_basis, x):
    """Values of all n_basis B-splines at one point x (clamped knot vector)."""
    t = knots
    x = float(np.clip(x, t[d], t[n_basis]))
    # span: t[span] <= x < t[span+1], with the right end folded into the last span
    span = int(np.searchsorted(t, x, side="right") - 1)
    span = min(max(span, d), n_basis - 1)
    P = span_polynomials(t, d, span, t[span])
    u = x - t[span]
    vals = np.zeros(n_basis)
    vals[span - d:span + 1] = P @ (u ** np.arange(d + 1))
    return vals


def bspline_slopes(knots, d, n_basis, x):
    """Arc-length derivatives of all n_basis B-splines at one point x: the
    derivative of the span's polynomial (the span chosen as bspline_values
    chooses it, right-continuous, the last span at the right end)."""
    t = knots
    x = float(np.clip(x, t[d], t[n_basis]))
    span = int(np.searchsorted(t, x, side="right") - 1)
    span = min(max(span, d), n_basis - 1)
    P = span_polynomials(t, d, span, t[span])
    u = x - t[span]
    powers = np.arange(d + 1)
    du = np.where(powers > 0, powers * u ** np.maximum(powers - 1, 0), 0.0)
    vals = np.zeros(n_basis)
    vals[span - d:span + 1] = P @ du
    return vals


# ---------------------------------------------------------------------------
# The mesh
# ---------------------------------------------------------------------------
@dataclass
class HostMesh:
    degree: int
    n_seg: int
    n_basis: int
    seg: np.ndarray          # (N, 12)
    seg_edge: np.ndarray     # (N, 2) int32
    edge: np.ndarray         # (E, 2)
    edge_i: np.ndarray       # (E, 2) int32
    se_edge: np.ndarray      # (n_se,) int32
    band_poly: np.ndarray    # (N, 3, 3)
    band_basis: np.ndarray   # (N, 3) int32, -1 empty
    supp: np.ndarray         # (n_b, 3) int32, -1 absent wing
    polys: np.ndarray        # (n_b, 3, 3)
    bw_slot: np.ndarray      # (n_b, 3) int32
    bw_loc: np.ndarray       # (n_b, 3) int32
    tile_b0: np.ndarray      # (T+1,) int32
    tile_seg: np.ndarray     # (T, TSMAX) int32
    tile_nseg: np.ndarray    # (T,) int32
    tile_pairs: np.ndarray   # (P, 2) int32, ti <= tj
    symmetric: bool
    gl: np.ndarray           # (2*n_qp + 32,)
    n_qp: int                # Gauss-Legendre points per segment per axis
    per_wire: list = field(default_factory=list)  # arc, knots, kept, l2g, seg_off
    kcl: list = field(default_factory=list)       # [(junction, [(m, sign)])]


def build_mesh(wires, npe, radius, degree, junctions, n_qp=4, ground_ends=()):
    """Discretise the wires and build every k-independent table.

    ground_ends: (wire, "start"|"end") ends lying in an active ground plane
    with no junction there; each keeps its value-1 boundary basis (kind
    "gnd", no KCL partner: the image carries the current on through the
    plane; reference bspline.py:1108-1151, 1296-1316)."""
    d = int(degree)
    n_w = len(wires)
    st = ["free"] * n_w
    en = ["free"] * n_w
    for j, jw in enumerate(junctions):
        for w, end in jw:
            if end == "start":
                st[w] = j
            else:
                en[w] = j
    for w, end in ground_ends:
        if end == "start" and st[w] == "free":
            st[w] = "ground"
        elif end == "end" and en[w] == "free":
            en[w] = "ground"

    seg_rows, seg_edge, edge_rows, edge_i, se_edge = [], [], [], [], []
    per_wire = []
    n_seg = 0
    n_se = 0
    eid = 0
    for w, pl in enumerate(wires):
        pl = np.asarray(pl, dtype=np.float64)
        seg_off = n_seg
        hs = []
        for e in range(pl.shape[0] - 1):
            p0, p1 = pl[e], pl[e + 1]
            ev = p1 - p0
            L = float(np.linalg.norm(ev))
            if L < 1e-15:
                raise ValueError(f"wire {w} edge {e} has zero length")
            tan = ev / L
            n_e = int(npe[w][e])
            t = np.linspace(0.0, 1.0, n_e + 1)
            pts = (1 - t[:, None]) * p0[None, :] + t[:, None] * p1[None, :]
            for s in range(n_e):
                l, r = pts[s], pts[s + 1]
                seg_rows.append(np.concatenate([l, r, tan, [np.linalg.norm(r - l), radius[w], 0.0]]))
                seg_edge.append((eid, s))
            arc = np.linspace(0.0, L, n_e + 1)
            edge_rows.append((arc[1] - arc[0], radius[w]))
            edge_i.append((n_e, n_se))
            se_edge.extend([eid] * (2 * n_e - 1))
            n_se += 2 * n_e - 1
            hs.append(np.full(n_e, L / n_e))
            n_seg += n_e
            eid += 1
        h = np.concatenate(hs)
        arc = np.concatenate([[0.0], np.cumsum(h)])
        per_wire.append(dict(seg_off=seg_off, n=len(h), h=h, arc=arc))

    # bases
    # The device tables hold NM = 3 wings / band slots / powers per basis and
    # segment; a degree-1 (tent) basis fills two of each, the rest stays empty
    # (wing -1, zero coefficients).
    NM = 3
    if not 1 <= d <= 2:
        raise ValueError(f"degree {d} is not built (1 or 2)")
    supp, polys, slots = [], [], []
    band_poly = np.zeros((n_seg, NM, NM))
    band_basis = np.full((n_seg, NM), -1, dtype=np.int32)
    kcl_dirs = {j: [] for j in range(len(junctions))}
    m = 0
    for w, pw in enumerate(per_wire):
        arc, n_sw = pw["arc"], pw["n"]
        knots = np.concatenate([np.zeros(d), arc, np.full(d, arc[-1])])
        nb_w = n_sw + d
        span_p = [span_polynomials(knots, d, s + d, knots[s + d]) for s in range(n_sw)]
        kept = []
        if st[w] == "ground":
            kept.append((0, "gnd", None, "start"))
        elif st[w] != "free":
            kept.append((0, "dir", st[w], "start"))
        kept += [(j, "int", None, None) for j in range(1, nb_w - 1)]
        if en[w] == "ground":
            kept.append((nb_w - 1, "gnd", None, "end"))
        elif en[w] != "free":
            kept.append((nb_w - 1, "dir", en[w], "end"))
        l2g = {}
        for kidx, (j, kind, jidx, endp) in enumerate(kept):
            lo, hi = max(0, j - d), min(n_sw, j + 1)
            s_m = np.full(NM, -1, dtype=np.int32)
            p_m = np.zeros((NM, NM))
            sl_m = np.full(NM, -1, dtype=np.int32)
            for a, s in enumerate(range(lo, hi)):
                slot = j - s  # row of basis j among the span's bases s..s+d
                s_m[a] = pw["seg_off"] + s
                p_m[a, :d + 1] = span_p[s][slot]
                sl_m[a] = slot
                band_poly[pw["seg_off"] + s, slot] = p_m[a]
                band_basis[pw["seg_off"] + s, slot] = m
            supp.append(s_m)
            polys.append(p_m)
            slots.append(sl_m)
            l2g[kidx] = m
            if kind == "dir":
                kcl_dirs[jidx].append((m, 1.0 if endp == "start" else -1.0))
            m += 1
        pw.update(knots=knots, kept=kept, l2g=l2g, nb_w=nb_w)
    n_b = m
    supp = np.array(supp, dtype=np.int32).reshape(n_b, NM)
    polys = np.array(polys).reshape(n_b, NM, NM)
    slots = np.array(slots, dtype=np.int32).reshape(n_b, NM)

    # basis tiles: consecutive bases, <= TB per tile, <= TSMAX distinct segments
    tile_b0, tile_segs = [0], []
    cur = set()
    for mm in range(n_b):
        segs = set(int(s) for s in supp[mm] if s >= 0)
        if mm - tile_b0[-1] >= TB or len(cur | segs) > TSMAX:
            tile_segs.append(sorted(cur))
            tile_b0.append(mm)
            cur = set()
        cur |= segs
    tile_segs.append(sorted(cur))
    tile_b0.append(n_b)
    n_t = len(tile_segs)
    tile_seg = np.full((n_t, TSMAX), -1, dtype=np.int32)
    tile_nseg = np.zeros(n_t, dtype=np.int32)
    bw_loc = np.full((n_b, NM), -1, dtype=np.int32)
    for t_i, segs in enumerate(tile_segs):
        tile_seg[t_i, :len(segs)] = segs
        tile_nseg[t_i] = len(segs)
        pos = {s: i for i, s in enumerate(segs)}
        for mm in range(tile_b0[t_i], tile_b0[t_i + 1]):
            for a in range(NM):
                if supp[mm, a] >= 0:
                    bw_loc[mm, a] = pos[int(supp[mm, a])]
    ti, tj = np.triu_indices(n_t)
    tile_pairs = np.stack([ti, tj], axis=1).astype(np.int32)

    xi, wi = np.polynomial.legendre.leggauss(n_qp)
    xf, wf = np.polynomial.legendre.leggauss(16)
    gl = np.concatenate([0.5 * (xi + 1.0), 0.5 * wi, xf, wf])

    kcl = [(j, kcl_dirs[j]) for j in range(len(junctions))]
    return HostMesh(
        degree=d, n_seg=n_seg, n_basis=n_b,
        seg=np.array(seg_rows), seg_edge=np.array(seg_edge, dtype=np.int32),
        edge=np.array(edge_rows), edge_i=np.array(edge_i, dtype=np.int32),
        se_edge=np.array(se_edge, dtype=np.int32),
        band_poly=band_poly, band_basis=band_basis, supp=supp, polys=polys,
        bw_slot=slots, bw_loc=bw_loc, tile_b0=np.array(tile_b0, dtype=np.int32),
        tile_seg=tile_seg, tile_nseg=tile_nseg, tile_pairs=tile_pairs,
        symmetric=bool(np.all(np.asarray(radius) == radius[0])), gl=gl, n_qp=int(n_qp),
        per_wire=per_wire, kcl=kcl)


def coaxial_labels(seg_l, seg_r, tangents, seg_a, tol=1e-6):
    """Coaxial-and-equal-radius labels of segments, for the extended
    thin-wire kernel: a pair is extended iff its labels agree.  A segment
    joins the first earlier representative with |t . t_rep| >= 1 - tol,
    |a / a_rep - 1| <= tol and its centre within tol * a_rep of the
    representative's line (perpendicular to its own axis), else it starts a
    label (the rule and thresholds of reference
    _bspline_kernels.py:249-322).  O(N G) for G labels."""
    seg_l = np.asarray(seg_l, dtype=np.float64)
    tangents = np.asarray(tangents, dtype=np.float64)
    seg_a = np.asarray(seg_a, dtype=np.float64)
    centers = 0.5 * (seg_l + np.asarray(seg_r, dtype=np.float64))
    labels = np.empty(len(seg_a), dtype=np.int32)
    reps = []
    for i in range(len(seg_a)):
        label = len(reps)
        for g, r in enumerate(reps):
            if abs(tangents[r] @ tangents[i]) < 1.0 - tol or abs(seg_a[i] / seg_a[r] - 1.0) > tol:
                continue
            d = centers[i] - centers[r]
            if np.linalg.norm(d - (d @ tangents[i]) * tangents[i]) <= tol * seg_a[r]:
                label = g
                break
        if label == len(reps):
            reps.append(i)
        labels[i] = label
    return labels


def ek_tables(mesh, ground_z=None):
    """(labels, image observer labels, image source labels) of a mesh for the
    extended kernel; the image pair comes from one scan over the real and the
    mirrored segments so both sides stay comparable (None without ground)."""
    seg = mesh.seg
    sl, sr, t, a = seg[:, 0:3], seg[:, 3:6], seg[:, 6:9], seg[:, 10]
    lab = coaxial_labels(sl, sr, t, a)
    if ground_z is None:
        return lab, None, None
    mir = np.array([1.0, 1.0, -1.0])
    off = np.array([0.0, 0.0, 2.0 * ground_z])
    joint = coaxial_labels(np.vstack([sl, sl * mir + off]), np.vstack([sr, sr * mir + off]),
                           np.vstack([t, t * mir]), np.concatenate([a, a]))
    n = len(a)
    return lab, joint[:n], joint[n:]


# ---------------------------------------------------------------------------
# Sparse port columns and constraint rows
# ---------------------------------------------------------------------------
def sparse_pack(cols):
    """[(idx list, val list)] -> padded (len, maxnz) int32 / float64 arrays."""
    maxnz = max([len(i) for i, _v in cols] + [1])
    idx = np.full((len(cols), maxnz), -1, dtype=np.int32)
    val = np.zeros((len(cols), maxnz))
    for c, (i, v) in enumerate(cols):
        idx[c, :len(i)] = i
        val[c, :len(v)] = v
    return idx, val, maxnz


def feed_column(mesh, wire, s_f, feed_model="point", smoothing=None, n_qp_source=16):
    """The drive (and readout) column of a gap feed at arclength s_f of a wire
    (reference bspline.py:2806-2942):

    * point delta gap: v_m = Phi_m(s_f);
    * segment gap (feed_model="segment"): v_m = (1/D) int Phi_m over the mesh
      cell [s_lo, s_hi] holding s_f (a feed on a knot takes the cell to its
      right, one at the wire end the last cell), D = s_hi - s_lo;
    * smoothed (smoothing = alpha): v_m = int Phi_m g_w(s - s_f) ds with the
      cos^2 bump g_w(x) = (2/w) cos^2(pi x / w) on |x| < w/2, w = alpha h of
      the cell holding s_f, clipped to the wire.

    The spread forms integrate with n_qp_source-point Gauss-Legendre over the
    cell or the bump's support."""
    pw = mesh.per_wire[wire]
    d = mesh.degree
    arc = np.asarray(pw["arc"], dtype=float)
    s_f = arc[-1] / 2.0 if s_f is None else float(s_f)
    if smoothing is None and feed_model == "point":
        v_full = bspline_values(pw["knots"], d, pw["nb_w"], s_f)
    else:
        cell = int(np.searchsorted(arc, s_f, side="right")) - 1
        cell = max(0, min(cell, len(arc) - 2))
        xi, wq = np.polynomial.legendre.leggauss(int(n_qp_source))
        if feed_model == "segment":
            s_lo, s_hi = float(arc[cell]), float(arc[cell + 1])
            h = s_hi - s_lo
            t = 0.5 * (s_hi + s_lo) + 0.5 * h * xi
            g = (0.5 * h * wq) / h
        else:
            width = float(smoothing) * float(arc[cell + 1] - arc[cell])
            half = width / 2.0
            s_lo, s_hi = max(0.0, s_f - half), min(float(arc[-1]), s_f + half)
            if s_lo >= s_hi:
                raise ValueError("feed_smoothing_factor too large for wire -- bump doesn't fit")
            t = 0.5 * (s_hi + s_lo) + 0.5 * (s_hi - s_lo) * xi
            dx = t - s_f
            bump = np.where(np.abs(dx) < half, (2.0 / width) * np.cos(np.pi * dx / width) ** 2, 0.0)
            g = bump * (0.5 * (s_hi - s_lo) * wq)
        vals = np.array([bspline_values(pw["knots"], d, pw["nb_w"], x) for x in t])
        v_full = vals.T @ g
    idx, val = [], []
    for kidx, (j, *_r) in enumerate(pw["kept"]):
        if v_full[j] != 0.0:
            idx.append(pw["l2g"][kidx])
            val.append(float(v_full[j]))
    return idx, val
