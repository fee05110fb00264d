"""The deck front door: NEC decks on the MI355X path through the
reference's own deck front end.

The reference turns a deck into a solver in two verbs --
``momwire.deck.parse(text)`` and ``build_solver(model, basis=...)`` -- and
``build_solver`` picks the solver class from its roster ``BASES``
(reference src/momwire/deck/_solver.py:84-112, 472-606): one constructor
call with ``wires``, ``n_per_edge_per_wire``, ``feeds``, ``wavelength``,
``wire_radius``, ``junctions``, ``node_gaps``, ``extended_kernel``, the
wire-loading arrays and the ground keywords.  This module's
``BSplineSolver`` takes every one of those keywords with the reference's
meaning, so the GPU path joins the roster as one more basis:

    from momwire import deck
    from momwire_amd import deck as gpu_deck
    gpu_deck.register(deck._solver)                 # adds "bspline-gpu"
    built = deck.build_solver(deck.parse(text), basis="bspline-gpu")
    y = built.solver.compute_port_solution().y      # on the GPU

The deck grammar, the connection snapping, the chaining into polylines and
the port plan stay the reference's own code, untouched: what a deck MEANS is
decided in one place, and the GPU path answers the same constructor call.
``solver_from_kwargs`` rebuilds a solver from recorded constructor keywords
(complex values as ``{"complex": [re, im]}``), which is how the committed
deck fixtures replay on a machine without the reference installed.
"""

from __future__ import annotations

from types import MappingProxyType

from .bspline import BSplineSolver

GPU_BASIS = "bspline-gpu"


def register(solver_module, name: str = GPU_BASIS, replace_default: bool = False):
    """Add the GPU solver to a deck front end's basis roster.

    ``solver_module`` is the module holding the roster (the reference's
    ``momwire.deck._solver``); its ``BASES`` mapping is replaced by a copy
    with ``name`` -> (this package's BSplineSolver, no extra keywords), and,
    with ``replace_default``, the ``"bspline"`` entry pointed at it too.
    Returns the new roster."""
    roster = dict(solver_module.BASES)
    roster[name] = (BSplineSolver, MappingProxyType({}))
    if replace_default:
        roster["bspline"] = roster[name]
    solver_module.BASES = MappingProxyType(roster)
    return solver_module.BASES


def _decode(v):
    if isinstance(v, dict) and set(v) == {"complex"}:
        re, im = v["complex"]
        return complex(re, im)
    if isinstance(v, list):
        return [_decode(x) for x in v]
    if isinstance(v, dict):
        return {k: _decode(x) for k, x in v.items()}
    return v


def solver_from_kwargs(kwargs, **overrides):
    """A GPU BSplineSolver from recorded deck constructor keywords."""
    kw = _decode(dict(kwargs))
    kw.update(overrides)
    return BSplineSolver(**kw)
