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
