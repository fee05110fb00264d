"""Synthetic structures the port-solution path is measured and tested on.

Every structure is a plain keyword dict for ``BSplineSolver(**kw)``; seeds are
fixed so the same geometry is rebuilt on every machine.

* ``straight_dipole``  -- one 21-segment wire, (-5.05,0,10)-(5.05,0,10) m,
  radius 1 mm, 14.1 MHz, centre feed (the reference README's deck geometry).
* ``dipole_array``     -- an nx x ny planar array of z-directed 0.48-lambda
  dipoles at 0.6-lambda pitch with seeded +-0.02-lambda x/y jitter, one centre
  feed per dipole.  8x8 x 41 segments is 2,624 unknowns and 64 ports;
  32x32 x 15 segments is 15,360 unknowns and 1,024 ports.
* ``junction_tee``     -- a small three-wire junction with a bent arm, used to
  exercise the KCL constraint rows, junction ports and node gaps.
* ``random_structure`` -- seeded random polyline structures (1-12 wires sharing
  junction nodes) for the batched sweep.
* ``grounded_array``   -- an nx x ny array of horizontal x-directed 0.48-lambda
  dipoles at 0.5-lambda pitch over lossy Sommerfeld ground; height, eps_r,
  sigma and frequency are drawn from a seeded generator (4x4 x 21 segments is
  336 unknowns and 16 ports).
* ``grounded_tee``     -- the junction tee above Sommerfeld ground (vertical,
  horizontal and slanted segments, so every remainder surface is exercised).
"""

import numpy as np

C_LIGHT = 299792458.0


def straight_dipole(n_seg=21, freq_hz=14.1e6):
    return dict(
        wires=[[(-5.05, 0.0, 10.0), (5.05, 0.0, 10.0)]],
        n_per_edge_per_wire=[[n_seg]],
        degree=2,
        wire_radius=1e-3,
        wavelength=C_LIGHT / freq_hz,
        feeds=[(0, None, 1.0)],
    )


def dipole_array(nx=8, ny=8, n_seg=41, seed=1, freq_hz=300e6, length=0.48,
                 radius=1e-3, pitch=0.6, jitter=0.02):
    lam = C_LIGHT / freq_hz
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(nx * ny, 2)) * lam
    half = 0.5 * length * lam
    wires = []
    for ix in range(nx):
        for iy in range(ny):
            e = ix * ny + iy
            x = ix * pitch * lam + jit[e, 0]
            y = iy * pitch * lam + jit[e, 1]
            wires.append([(x, y, -half), (x, y, half)])
    return dict(
        wires=wires,
        n_per_edge_per_wire=[[n_seg]] * len(wires),
        degree=2,
        wire_radius=radius * lam,
        wavelength=lam,
        feeds=[(w, None, 1.0) for w in range(len(wires))],
    )


def junction_tee(ports="feed"):
    """Vertical mast, two horizontal arms (one bent) meeting at its top."""
    wires = [
        [(0.0, 0.0, 0.05), (0.0, 0.0, 0.45)],
        [(0.0, 0.0, 0.45), (0.20, 0.0, 0.45), (0.30, 0.10, 0.40)],
        [(0.0, 0.0, 0.45), (-0.25, 0.0, 0.45)],
    ]
    kw = dict(
        wires=wires,
        n_per_edge_per_wire=[[9], [7, 4], [8]],
        degree=2,
        wire_radius=1e-3,
        wavelength=1.0,
        junctions=[[(0, "end"), (1, "start"), (2, "start")]],
        feeds=[(0, None, 1.0)],
    )
    if ports == "junction_port":
        kw["feeds"] = [(0, None, 1.0), (2, 0.1, 1.0)]
        kw["junction_ports"] = [0]
    elif ports == "node_gap":
        kw["feeds"] = [(0, None, 1.0)]
        kw["node_gaps"] = [(1, "start", 1.0)]
    return kw


def wire_chain(n_wires=40, n_seg=3, feed_every=None, wavelength=1.0):
    """A zig-zag of n_wires straight wires joined end to start (n_wires - 1
    two-wire junctions, each a KCL constraint row): more constraint rows
    than the 32 the small Schur solve takes.  feed_every=m puts a centre
    feed on every m-th wire (else one feed on wire 0)."""
    pts = [(0.0, 0.0, 0.0)]
    for i in range(n_wires):
        x, _y, z = pts[-1]
        pts.append((x + 0.03, 0.012 * (1 if i % 2 == 0 else -1), z + 0.004 * (i % 3)))
    wires = [[pts[i], pts[i + 1]] for i in range(n_wires)]
    junctions = [[(i, "end"), (i + 1, "start")] for i in range(n_wires - 1)]
    feeds = ([(i, None, 1.0) for i in range(0, n_wires, feed_every)] if feed_every
             else [(0, None, 1.0)])
    return dict(wires=wires, n_per_edge_per_wire=[[n_seg]] * n_wires, degree=2,
                wire_radius=5e-4, wavelength=wavelength, junctions=junctions, feeds=feeds)


def grounded_monopole(kind="pec", base_feed=True):
    """A quarter-wave vertical standing on the ground plane (ground contact:
    the base keeps its value-1 basis, the image carries the current on):
    base-fed (s = 0) or fed at the first segment's centre, over PEC or the
    Sommerfeld ground (eps_r 13, sigma 0.005 S/m) at 14 MHz; 'radials' adds
    three sloping radials meeting the vertical at a junction in the plane
    (a grounded junction: its KCL row is dropped)."""
    lam = 299792458.0 / 14e6
    h = 0.25 * lam
    n = 17
    kw = dict(wires=[[(0.0, 0.0, 0.0), (0.0, 0.0, h)]], n_per_edge_per_wire=[[n]], degree=2,
              wire_radius=0.01, wavelength=lam, ground_z=0.0,
              feeds=[(0, 0.0 if base_feed else 0.5 * h / n, 1.0)])
    if kind == "sommerfeld":
        kw.update(ground_eps=(13.0, 0.005), ground_model="sommerfeld")
    if kind == "radials":
        kw["wires"] = [[(0.0, 0.0, 0.0), (0.0, 0.0, h)]] + [
            [(0.0, 0.0, 0.0), (0.2 * lam * np.cos(t), 0.2 * lam * np.sin(t), 0.05 * lam)]
            for t in (0.3, 2.4, 4.5)]
        kw["n_per_edge_per_wire"] = [[n], [9], [9], [9]]
        kw["junctions"] = [[(0, "start"), (1, "start"), (2, "start"), (3, "start")]]
        kw["feeds"] = [(0, 0.5 * h / n, 1.0)]
    return kw


def ground_parameters(seed):
    """(height / lambda, eps_r, sigma S/m, frequency Hz) of one grounded-array
    draw: height U(0.05, 1) lambda, eps_r U(5, 30), sigma log-U(1e-3, 5e-2),
    frequency U(3, 30) MHz."""
    rng = np.random.default_rng(seed)
    height = float(rng.uniform(0.05, 1.0))
    eps_r = float(rng.uniform(5.0, 30.0))
    sigma = float(np.exp(rng.uniform(np.log(1e-3), np.log(5e-2))))
    freq = float(rng.uniform(3e6, 30e6))
    return height, eps_r, sigma, freq


def ground_box_draws(n=16, seed=2026):
    """n grounded-array draws spread over the whole C5 box: heights stratified
    over [0.05, 1] wavelength (one per n-th of the range, jittered), eps_r
    U(5, 30), sigma log-U(1e-3, 5e-2) S/m, frequency U(3, 30) MHz; each a
    (height / lambda, eps_r, sigma, freq) tuple for grounded_array(params=...)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        h = 0.05 + 0.95 * (i + rng.uniform()) / n
        er = float(rng.uniform(5.0, 30.0))
        sig = float(np.exp(rng.uniform(np.log(1e-3), np.log(5e-2))))
        f = float(rng.uniform(3e6, 30e6))
        out.append((float(h), er, sig, f))
    return out


def grounded_array(nx=4, ny=4, n_seg=21, seed=4, length=0.48, radius=1e-3, pitch=0.5,
                   params=None, ground_model="sommerfeld", phi_mode=None):
    """Horizontal dipole array over finite ground (ground plane z = 0),
    Sommerfeld by default; ``ground_model="refl-coef"`` (with an optional
    ``phi_mode``) for the reflection-coefficient ground.

    ``params`` overrides the seeded (height/lambda, eps_r, sigma, freq) draw.
    """
    height, eps_r, sigma, freq = params if params is not None else ground_parameters(seed)
    lam = C_LIGHT / freq
    half = 0.5 * length * lam
    z = height * lam
    wires = []
    for ix in range(nx):
        for iy in range(ny):
            x = ix * pitch * lam
            y = iy * pitch * lam
            wires.append([(x - half, y, z), (x + half, y, z)])
    return dict(
        wires=wires,
        n_per_edge_per_wire=[[n_seg]] * len(wires),
        degree=2,
        wire_radius=radius * lam,
        wavelength=lam,
        feeds=[(w, None, 1.0) for w in range(len(wires))],
        ground_z=0.0,
        ground_eps=(eps_r, sigma),
        ground_model=ground_model,
        **({} if phi_mode is None else {"ground_phi_mode": phi_mode}),
    )


def grounded_tee(eps_r=12.0, sigma=5e-3, freq_hz=10e6, ground_model="sommerfeld", phi_mode=None):
    """The junction tee scaled to ``freq_hz`` and lifted above finite ground
    (Sommerfeld by default)."""
    kw = junction_tee("feed")
    lam = C_LIGHT / freq_hz
    kw["wires"] = [[(p[0] * lam, p[1] * lam, p[2] * lam) for p in w] for w in kw["wires"]]
    kw["wire_radius"] = 1e-3 * lam
    kw["wavelength"] = lam
    kw.update(ground_z=0.0, ground_eps=(eps_r, sigma), ground_model=ground_model)
    if phi_mode is not None:
        kw["ground_phi_mode"] = phi_mode
    return kw


def _seg_seg_distance(p0, p1, q0, q1):
    """Minimum distance between segments p0-p1 and q0-q1 (sampled exactly
    enough for a clearance test: endpoints against segments plus the
    closest-approach of the two carrier lines when it falls inside both)."""
    def pt_seg(p, a, b):
        ab = b - a
        t = np.clip(np.dot(p - a, ab) / max(np.dot(ab, ab), 1e-300), 0.0, 1.0)
        return np.linalg.norm(p - (a + t * ab))
    d = min(pt_seg(p0, q0, q1), pt_seg(p1, q0, q1), pt_seg(q0, p0, p1), pt_seg(q1, p0, p1))
    u, v, w = p1 - p0, q1 - q0, p0 - q0
    a, b, c, dd, e = u @ u, u @ v, v @ v, u @ w, v @ w
    den = a * c - b * b
    if den > 1e-12 * a * c:
        s = (b * e - c * dd) / den
        t = (a * e - b * dd) / den
        if 0 < s < 1 and 0 < t < 1:
            d = min(d, np.linalg.norm(w + s * u - t * v))
    return d


def random_structure(seed, f0_hz=300e6, nb_range=(40, 2000), max_tries=200):
    """One seeded random wire structure for the batched sweep.

    1-12 polylines of 1-3 straight edges each, grown as a tree: every wire
    after the first starts at an existing node (a junction shared with the
    wires already there).  Each wire's length is U(0.5, 6) lambda0, split over
    its edges; radius log-U(1e-4, 1e-2) lambda0; one segment length
    U(lambda0/20, lambda0/10) per structure; 1-4 delta-gap feeds at random
    interior arc positions.  Edges keep a clearance of one segment length
    from every edge they do not touch.  Structures whose unknown count falls
    outside nb_range are redrawn.
    """
    lam = C_LIGHT / f0_hz
    rng = np.random.default_rng(seed)
    for _ in range(max_tries):
        n_w = int(rng.integers(1, 13))
        seg_len = rng.uniform(lam / 20, lam / 10)
        radius = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))) * lam)
        nodes = [np.zeros(3)]
        wires, npe, ends = [], [], []  # ends: (start node id, end node id)
        edges = []  # (p0, p1, wire)
        ok = True
        for w in range(n_w):
            start = int(rng.integers(0, len(nodes))) if w else 0
            n_e = int(rng.integers(1, 4))
            length = rng.uniform(0.5, 6.0) * lam
            parts = rng.dirichlet(np.ones(n_e)) * length
            pts = [nodes[start]]
            direc = rng.normal(size=3)
            placed = True
            for e in range(n_e):
                for _try in range(30):
                    d = direc + 0.8 * rng.normal(size=3)
                    d /= np.linalg.norm(d)
                    p1 = pts[-1] + max(parts[e], 2 * seg_len) * d
                    clear = True
                    for (a, b, ww) in edges:
                        touching = (np.linalg.norm(a - pts[-1]) < 1e-9 or
                                    np.linalg.norm(b - pts[-1]) < 1e-9)
                        if touching:
                            # share only the junction point: keep an angle
                            other = b if np.linalg.norm(a - pts[-1]) < 1e-9 else a
                            cosang = np.dot(other - pts[-1], d) / np.linalg.norm(other - pts[-1])
                            if cosang > 0.9:
                                clear = False
                                break
                            continue
                        if _seg_seg_distance(pts[-1], p1, a, b) < seg_len:
                            clear = False
                            break
                    if clear:
                        break
                else:
                    placed = False
                    break
                edges.append((pts[-1], p1, w))
                pts.append(p1)
                direc = d
            if not placed:
                ok = False
                break
            nodes.append(pts[-1])
            wires.append([tuple(p) for p in pts])
            npe.append([max(2, int(round(np.linalg.norm(pts[i + 1] - pts[i]) / seg_len)))
                        for i in range(len(pts) - 1)])
            ends.append((start, len(nodes) - 1))
        if not ok:
            continue
        # junction groups: every node touched by >= 2 wire ends
        groups = {}
        for w, (s_node, e_node) in enumerate(ends):
            groups.setdefault(s_node, []).append((w, "start"))
            groups.setdefault(e_node, []).append((w, "end"))
        junctions = [g for g in groups.values() if len(g) >= 2]
        n_junc_ends = sum(len(g) for g in junctions)
        n_b = sum(sum(e) for e in npe) + 2 * n_w * 0 + n_junc_ends - 0
        # unknowns: per wire N + d - 2 interior bases plus kept junction-end bases
        n_b = sum(sum(e) for e in npe) + n_junc_ends
        if not (nb_range[0] <= n_b <= nb_range[1]):
            continue
        n_feeds = int(rng.integers(1, 5))
        feed_w = rng.integers(0, n_w, size=n_feeds)
        feeds = []
        for w in feed_w:
            L = sum(np.linalg.norm(np.subtract(wires[w][i + 1], wires[w][i]))
                    for i in range(len(wires[w]) - 1))
            feeds.append((int(w), float(rng.uniform(0.2, 0.8) * L), 1.0))
        return dict(wires=wires, n_per_edge_per_wire=npe, degree=2, wire_radius=radius,
                    wavelength=lam, junctions=junctions, feeds=feeds)
    raise RuntimeError(f"no admissible structure for seed {seed}")


def sweep_structures(n=128, seed=3):
    """The batched-sweep workload: n seeded random structures."""
    return [random_structure(seed * 100003 + i) for i in range(n)]


def sweep_wavenumbers(n_freq=16, f0_hz=300e6, lo=0.8, hi=1.2):
    return 2 * np.pi * np.linspace(lo, hi, n_freq) * f0_hz / C_LIGHT


def ek_structure(name):
    """Structures for the extended thin-wire kernel (``extended_kernel=True``,
    reference bspline.py:393-450): fat wires (radius a few 1e-3 lambda) so
    the kernel's O(a^2) factor shows.

    dipole         one straight wire (every pair coaxial)
    collinear      two collinear wires meeting at a junction (coaxial
                   pairs across wires)
    tee            the junction tee (a bend and a K = 3 junction: mostly
                   non-coaxial pairs)
    pec_vertical   a vertical wire and a horizontal one over PEC ground (the
                   vertical wire's image is coaxial with it)
    refl_dipole    a horizontal dipole over the reflection-coefficient ground
    somm_tee       the tee over Sommerfeld ground (remainder kept reduced)
    """
    if name == "dipole":
        return dict(wires=[[(0.0, 0.0, -0.24), (0.0, 0.0, 0.24)]], n_per_edge_per_wire=[[21]],
                    degree=2, wire_radius=4e-3, wavelength=1.0, feeds=[(0, None, 1.0)],
                    extended_kernel=True)
    if name == "collinear":
        return dict(wires=[[(0.0, 0.0, -0.25), (0.0, 0.0, 0.0)], [(0.0, 0.0, 0.0), (0.0, 0.0, 0.25)]],
                    n_per_edge_per_wire=[[11], [11]], degree=2, wire_radius=3e-3, wavelength=1.0,
                    junctions=[[(0, "end"), (1, "start")]], feeds=[(0, 0.2, 1.0)],
                    extended_kernel=True)
    if name == "tee":
        kw = junction_tee("feed")
        kw.update(wire_radius=4e-3, extended_kernel=True)
        return kw
    if name == "pec_vertical":
        return dict(wires=[[(0.0, 0.0, 0.02), (0.0, 0.0, 0.27)], [(0.1, 0.0, 0.3), (0.4, 0.0, 0.3)]],
                    n_per_edge_per_wire=[[15], [13]], degree=2, wire_radius=4e-3, wavelength=1.0,
                    feeds=[(0, None, 1.0), (1, None, 1.0)], ground_z=0.0, extended_kernel=True)
    if name == "refl_dipole":
        kw = grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6), radius=4e-3,
                            ground_model="refl-coef")
        kw["extended_kernel"] = True
        return kw
    if name == "somm_tee":
        kw = grounded_tee()
        kw.update(wire_radius=4e-3 * kw["wavelength"], extended_kernel=True)
        return kw
    raise ValueError(f"unknown extended-kernel structure {name!r}")
