"""Attenuation accuracy against the reference's published figure (needs an MI355X).

For each of the 15 (alpha0, gamma) panels of the reference's ``figs/attenuation_modeling.svg``
(digitized into tests/golden/attenuation_modeling_reference.json), the engine propagates a plane
wave through the medium that ``Medium.build`` makes for (alpha0, gamma), and the attenuation is
measured by spectral ratio between two depths (tests/attenuation.py). At every frequency of the
figure the relative error against alpha0 f^gamma must stay within the reference's own error at
that panel and frequency, plus 2 % (1 % digitization, 1 % measurement).

The figure stops at alpha0 = 0.75 dB/(MHz^gamma cm); the heterogeneous configurations draw alpha0
up to 1.5. Panels at alpha0 = 1.0 and 1.5 (gamma 0.8, 1.1, 1.6) are held to the error envelope of
the nearest reference panel (alpha0 0.75, the nearest gamma), with the same slack. The reference's
error at another alpha0 only shows how large its fit error runs in that part of the band, not
where its sign changes, so the envelope at each frequency is the largest reference error at that
frequency and its two neighbours (it does not collapse where the reference curve crosses zero).
"""

import json
import os
from pathlib import Path

import numpy as np
import pytest

from fullwave25_amd import engine
from tests import attenuation as A

pytestmark = pytest.mark.gpu
SLACK = 0.02


def _relative_error(a0, g, f):
    """Measured attenuation / alpha0 f^gamma - 1 at the frequencies f (MHz)."""
    inp, meta = A.build_inputs(a0, g)
    print(f"alpha0 {a0} gamma {g}: grid {inp.shape}, {inp.nt} steps", flush=True)
    with engine.HipSimulation(inp) as sim:
        sim.run(inp.nt)
        traces = sim.sensors().astype(np.float64)
    return A.attenuation_from_traces(traces, meta, f) / A.target(a0, g, f) - 1


@pytest.mark.parametrize("panel", A.reference_panels(), ids=lambda p: f"a{p['alpha0_db_mhz_cm']}_g{p['gamma']}")
def test_attenuation_within_reference_error(panel):
    a0, g = panel["alpha0_db_mhz_cm"], panel["gamma"]
    print(f"panel {panel['panel']}", flush=True)
    f = np.array(panel["freq_mhz"])
    err = _relative_error(a0, g, f)
    ref = np.array(panel["rel_err"])
    print("  ours %:", np.round(100 * err, 2).tolist(), flush=True)
    print("  ref  %:", np.round(100 * ref, 2).tolist(), flush=True)
    out = Path(os.environ.get("FULLWAVE_TEST_OUT", ""))  # optional: keep the curves for inspection
    if os.environ.get("FULLWAVE_TEST_OUT") and out.is_dir():
        with open(out / "attenuation.jsonl", "a") as fh:
            fh.write(json.dumps({"panel": panel["panel"], "alpha0": a0, "gamma": g, "freq_mhz": f.tolist(),
                                 "ours": err.tolist(), "reference": ref.tolist()}) + "\n")
    assert np.all(np.abs(err) <= np.abs(ref) + SLACK), (err, ref)


def _nearest_panel(a0, g):
    return min(A.reference_panels(), key=lambda p: (abs(p["alpha0_db_mhz_cm"] - a0), abs(p["gamma"] - g)))


@pytest.mark.parametrize("a0,g", [(a, g) for a in (1.0, 1.5) for g in (0.8, 1.1, 1.6)], ids=lambda v: str(v))
def test_attenuation_beyond_reference_panels(a0, g):
    panel = _nearest_panel(a0, g)
    assert panel["alpha0_db_mhz_cm"] == 0.75
    f = np.array(panel["freq_mhz"])
    err = _relative_error(a0, g, f)
    ref = np.array(panel["rel_err"])
    a = np.abs(ref)
    envelope = np.maximum(a, np.maximum(np.r_[a[1:], a[-1]], np.r_[a[0], a[:-1]]))
    print(f"  envelope of panel {panel['panel']} (alpha0 0.75, gamma {panel['gamma']})", flush=True)
    print("  ours %:", np.round(100 * err, 2).tolist(), flush=True)
    print("  env  %:", np.round(100 * envelope, 2).tolist(), flush=True)
    assert np.all(np.abs(err) <= envelope + SLACK), (err, envelope)
