"""GPU parity of the readout of solved port columns: element moments and
knot steps (mw_element_moments), far-field moments over free space, PEC and
a finite ground (mw_far_moments, FP64 MFMA) and near fields E and H with and
without the PEC image (mw_near_fields), every port column in one call,
against the reference's own values (tests/golden/readout_*.npz) and the
oracle.

Bars: moments and steps 1e-13 of their largest entry (three taps and one
mean per entry); far-field moments 1e-11 of the pattern's largest (sums of
10^2 - 10^3 phased terms on MFMA against numpy's einsum); near fields 1e-10
of the largest (the point 0.01 wavelength off a wire cancels node charges
to ~1e-3 of their size).
"""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import readout, workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle import readout_oracle as ro

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CASES = {"readout_tee": lambda: W.junction_tee("junction_port"),
         "readout_array2x2": lambda: W.dipole_array(2, 2, 11, seed=5)}


def _say(*a):
    print("[gpu-test]", *a, flush=True)


class _Finite:
    kind = "refl"
    eps_r = 13.0
    sigma = 0.005


GROUNDS = {"free": None, "pec": "pec", "finite": _Finite()}


@pytest.mark.parametrize("case", sorted(CASES))
def test_element_moments_match_reference(case):
    _say("element moments", case)
    d = golden(case)
    s = BSplineSolver(**CASES[case]())
    X = torch.from_numpy(d["coeffs"]).cuda()
    for sub in (1, 8):
        mom, delta = readout.element_moments(s.readout_mesh(sub), X)
        assert rel_err(mom.cpu().numpy(), d[f"moment{sub}"]) < 1e-13
        assert rel_err(delta.cpu().numpy(), d[f"delta{sub}"]) < 1e-13


@pytest.mark.parametrize("case", sorted(CASES))
def test_far_moments_match_reference(case):
    _say("far-field moments", case)
    d = golden(case)
    k, gz, f = float(d["k"]), float(d["ground_z"]), float(d["freq_hz"])
    for name, g in GROUNDS.items():
        mt, mp = readout.far_moments(d["mid1"], d["moment1"], k, d["theta"], d["phi"], g, gz, f)
        rt, rp = d[f"far_{name}_theta"], d[f"far_{name}_phi"]
        scale = max(np.abs(rt).max(), np.abs(rp).max())
        et, ep = np.abs(mt - rt).max() / scale, np.abs(mp - rp).max() / scale
        _say(f"  {name}: {et:.2e} {ep:.2e}")
        assert et < 1e-11 and ep < 1e-11
        # one column through the reference's own single-column form
        m1t, m1p = readout.far_moments(d["mid1"], d["moment1"][..., 1], k, d["theta"], d["phi"],
                                       g, gz, f)
        assert m1t.shape == rt.shape[:2]
        assert np.abs(m1t - rt[..., 1]).max() < 1e-11 * scale
        assert np.abs(m1p - rp[..., 1]).max() < 1e-11 * scale


@pytest.mark.parametrize("case", sorted(CASES))
def test_near_fields_match_reference(case):
    _say("near fields", case)
    d = golden(case)
    k, gz, a = float(d["k"]), float(d["ground_z"]), float(d["radius"])
    els = (d["mid8"], d["moment8"], d["nodes8"], d["delta8"])
    for magnetic, tag in ((False, "E"), (True, "H")):
        f0 = readout.element_fields(d["points"], els, k, a, magnetic)
        fi = readout.element_fields(d["points"], els, k, a, magnetic, image_z=gz)
        e0, ei = rel_err(f0, d[f"near_{tag}"]), rel_err(fi, d[f"near_{tag}_pec"])
        _say(f"  {tag}: {e0:.2e}, with PEC image {ei:.2e}")
        assert e0 < 1e-10 and ei < 1e-10


def test_solver_readout_end_to_end():
    """Solve on the GPU, then read every port column out on the device: the
    reference's patterns and near fields of its own solution."""
    _say("solver readout end to end")
    d = golden("readout_array2x2")
    s = BSplineSolver(**CASES["readout_array2x2"]())
    mt, mp = s.port_far_moments(d["theta"], d["phi"])
    scale = np.abs(d["far_free_theta"]).max()
    assert np.abs(mt - d["far_free_theta"]).max() < 1e-10 * scale
    assert np.abs(mp - d["far_free_phi"]).max() < 1e-10 * scale
    E = s.port_near_fields(d["points"])
    assert rel_err(E, d["near_E"]) < 1e-9
    # the same columns handed over from the host
    mt2, _mp2 = s.port_far_moments(d["theta"], d["phi"], coeffs=d["coeffs"])
    assert np.abs(mt2 - d["far_free_theta"]).max() < 1e-11 * scale


@pytest.mark.parametrize("ground", ["free", "pec", "finite"])
def test_far_moments_ragged_tiles_match_oracle(ground):
    """Many workgroup tiles with ragged edges: 2,624 elements (82 chunks),
    37 columns (3 column tiles, the last of 5), 23 x 11 directions (4 tiles,
    the last of 61), random moments, against the oracle column by column."""
    _say("far-field moments, ragged tiles", ground)
    rng = np.random.default_rng(7)
    n, p = 2624, 37
    mid = rng.uniform(-2.0, 2.0, (n, 3))
    mid[:, 2] = np.abs(mid[:, 2]) + 0.1
    mom = rng.standard_normal((n, 3, p)) + 1j * rng.standard_normal((n, 3, p))
    th = np.linspace(0.0, np.pi, 23)
    ph = np.linspace(0.0, 2 * np.pi, 11, endpoint=False)
    k, gz, f = 2 * np.pi, -0.05, 3e8
    mt, mp = readout.far_moments(mid, mom, k, th, ph, GROUNDS[ground], gz, f)
    for j in (0, 16, 36):
        rt, rp = ro.far_moments(mid, mom[..., j], k, th, ph, ground, gz, 13.0, 0.005, f)
        scale = max(np.abs(rt).max(), np.abs(rp).max())
        assert np.abs(mt[..., j] - rt).max() < 1e-11 * scale
        assert np.abs(mp[..., j] - rp).max() < 1e-11 * scale
