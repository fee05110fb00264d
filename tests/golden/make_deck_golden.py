"""Write the deck-path golden vectors: for every NEC-2 deck of the
reference's own fixture corpus (tests/fixtures/nec_portal/*.deck in the
reference checkout), what the reference's deck front end --
``momwire.deck.parse`` then ``build_solver(model, basis="bspline")``
(reference src/momwire/deck/__init__.py:127-137, deck/_solver.py:472-606) --
hands the B-spline solver's constructor, and the port admittance matrix
that solver returns.

Runs ONLY in the development container (the reference checkout is mounted at
/root/reference); the reference's constructor keywords are recorded by
substituting a recording subclass for the roster's B-spline entry.  Outputs:
``decks/kwargs.json`` (constructor keywords per deck; complex numbers as
[re, im]; decks the reference front end refuses are listed with its message)
and ``decks/y.npz`` (Y per deck).

    python tests/golden/make_deck_golden.py
"""

import glob
import json
import os
import sys
import time
from types import MappingProxyType

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference/src")

import momwire  # noqa: E402
from momwire import deck  # noqa: E402
from momwire.bspline import BSplineSolver  # noqa: E402
from momwire.deck import _solver  # noqa: E402

DECKS = "/root/reference/tests/fixtures/nec_portal"


def _jsonable(v):
    if isinstance(v, dict):
        return {k: _jsonable(x) for k, x in v.items()}
    if isinstance(v, np.ndarray):
        return _jsonable(v.tolist())
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    if isinstance(v, complex):
        return {"complex": [v.real, v.imag]}
    if isinstance(v, (np.floating, np.integer)):
        return v.item()
    return v


class _Recording(BSplineSolver):
    last = None

    def __init__(self, **kw):
        _Recording.last = kw
        super().__init__(**kw)


def main():
    assert not momwire.accelerated, "expected the reference's pure-Python path"
    roster = dict(_solver.BASES)
    roster["bspline"] = (_Recording, MappingProxyType({}))
    _solver.BASES = MappingProxyType(roster)
    out_kw, out_y, refused = {}, {}, {}
    for path in sorted(glob.glob(os.path.join(DECKS, "*.deck"))):
        name = os.path.splitext(os.path.basename(path))[0]
        text = open(path).read()
        try:
            built = deck.build_solver(deck.parse(text), basis="bspline")
        except Exception as e:  # the reference front end's own refusal
            refused[name] = f"{type(e).__name__}: {e}"
            print(f"{name}: refused ({refused[name][:80]})", flush=True)
            continue
        kw = {k: v for k, v in _Recording.last.items() if k != "cancel"}
        t0 = time.time()
        y = built.solver.compute_port_solution().y
        out_kw[name] = _jsonable(kw)
        out_y[name] = np.asarray(y)
        print(f"{name}: {np.asarray(y).shape} ports, {time.time() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "decks", "kwargs.json"), "w") as f:
        json.dump({"kwargs": out_kw, "refused": refused}, f, indent=0)
    np.savez_compressed(os.path.join(HERE, "decks", "y.npz"), **out_y)


if __name__ == "__main__":
    main()
