"""Host-side pieces of the adjoint path (no GPU)."""

import numpy as np

from fullwave25_amd.adjoint import chain_to_c_rho, with_medium
from tests.helpers import raw_inputs


def test_chain_rule_to_c_and_rho():
    # K = rho c^2: dJ/dc = dJ/dK 2 rho c, dJ/drho = dJ/drho|K + dJ/dK c^2
    gc, gr = chain_to_c_rho(np.array([2.0]), np.array([3.0]), np.array([5.0]), np.array([7.0]))
    assert gc[0] == 5.0 * 2 * 3.0 * 2.0 and gr[0] == 7.0 + 5.0 * 4.0


def test_with_medium_sets_k_and_rho_only():
    inp = raw_inputs((20, 20), nt=5)
    c = np.full(inp.shape, 1500.0)
    rho = np.full(inp.shape, 1100.0)
    out = with_medium(inp, c, rho)
    assert out.maps["K"].dtype == np.float32 and np.allclose(out.maps["K"], 1100.0 * 1500.0**2)
    assert np.all(out.maps["rho"] == np.float32(1100.0))
    assert out.maps["kappax"] is inp.maps["kappax"] and out.dcmap is inp.dcmap
