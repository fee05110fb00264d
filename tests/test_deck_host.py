"""The deck front door on the CPU: every NEC-2 deck of the reference's own
fixture corpus, as the reference's deck front end builds it
(tests/golden/decks/kwargs.json, recorded by tests/golden/make_deck_golden.py
from the reference's parse + build_solver), constructs this package's
BSplineSolver with the same keywords and port plan; the roster hook adds the
GPU basis without disturbing the others."""

import json
import os
from types import MappingProxyType, SimpleNamespace

import numpy as np
import pytest

from conftest import GOLDEN
from momwire_amd import deck as gpu_deck
from momwire_amd.bspline import BSplineSolver

RECORD = json.load(open(os.path.join(GOLDEN, "decks", "kwargs.json")))
NAMES = sorted(RECORD["kwargs"])


def test_corpus_recorded():
    assert len(NAMES) == 60 and not RECORD["refused"]


@pytest.mark.parametrize("name", NAMES)
def test_deck_constructs_on_gpu_solver(name):
    y = np.load(os.path.join(GOLDEN, "decks", "y.npz"))[name]
    s = gpu_deck.solver_from_kwargs(RECORD["kwargs"][name])
    cols, _con = s._port_plan()
    assert len(cols) == y.shape[0] == s._port_count()
    assert s._build().n_basis > 0


def test_register_adds_gpu_basis():
    class Other:
        pass

    mod = SimpleNamespace(BASES=MappingProxyType({"bspline": (Other, MappingProxyType({}))}))
    roster = gpu_deck.register(mod)
    assert roster["bspline-gpu"][0] is BSplineSolver and roster["bspline"][0] is Other
    gpu_deck.register(mod, replace_default=True)
    assert mod.BASES["bspline"][0] is BSplineSolver
