"""GPU parity for the boundary types the airfoil decks never reach and for
the reference's stop paths (decks in tests/bc_decks.py; the oracle is
pinned to the reference on the same decks bit for bit by
test_oracle_ref.py).

* extrapolation / symmetry / periodic / free-stream boundaries (types 10,
  11, 22, 46 of bc and rbc, src/bc.f; and of the SA and SST boundary
  routines): 15 steps from a perturbed state, residual history to 1e-10 and
  state per cell to 1e-10 (FP64: the device's division, square root and pow
  may differ from the host's in the last bit);
* monitor stop (src/step.f:332-392): the GPU reports OT2D_MSTOP at the
  reference's iteration, with the state of that iteration;
* norm stop (rsum > 1000, src/step.f:105-110): OT2D_NORM at the first
  iteration with the same norms.
"""
import numpy as np
import pytest

from tests import bc_decks as B
from tests import helpers as H
from turns2d_urans_git_amd import abi

pytestmark = pytest.mark.gpu


def _pair(deck, segs, init):
    g = B.box()
    o = H.oracle_case(g, deck, segs=segs)
    d = H.gpu_case(g, deck, segs=segs)
    q0 = init(o.get_state(0))
    o.set_state(0, q0)
    d.set_state(0, q0)
    return o, d


@pytest.mark.parametrize("name", sorted(B.BC_DECKS))
def test_bc_types_vs_oracle(name):
    deck, segs = B.BC_DECKS[name]
    o, d = _pair(deck, segs, B.perturbed)
    rc_o, h_o, n_o = o.iterate(15)
    rc_d, h_d, n_d = d.iterate(15)
    assert rc_o == rc_d == 0 and n_o == n_d == 15
    np.testing.assert_allclose(h_d, h_o, rtol=1e-10)
    H.assert_cellwise(d.get_state(0), o.get_state(0), 1e-10, f"q ({name})")
    if not deck.get("lamin", 0):
        H.assert_cellwise(d.get_field(0, abi.F_TURMU), o.get_field(0, abi.F_TURMU), 1e-10, f"turmu ({name})")


@pytest.mark.parametrize("name", sorted(B.MSTOP_CASES))
@pytest.mark.parametrize("how", ["step", "iterate"])
def test_monitor_stop_vs_oracle(name, how):
    asq, amp, at = B.MSTOP_CASES[name]
    o, d = _pair(B.EULER, B.CLOSED, lambda q: B.quiescent(q, asq, amp))
    n_o = 0
    for _ in range(30):
        rc_o, _, _ = o.step()
        n_o += 1
        if rc_o:
            break
    assert rc_o == abi.MSTOP and n_o == at
    if how == "step":
        n_d = 0
        for _ in range(30):
            rc_d, _, _ = d.step()
            n_d += 1
            if rc_d:
                break
    else:
        rc_d, _, n_d = d.iterate(30)
    assert rc_d == abi.MSTOP and n_d == at
    H.assert_cellwise(d.get_state(0), o.get_state(0), 1e-10, "q at the stop")


def test_norm_stop_vs_oracle():
    o, d = _pair(B.EULER, B.EXTP_SYM_INF, B.blown_up)
    rc_o, n_o, _ = o.step()
    rc_d, n_d, _ = d.step()
    assert rc_o == rc_d == abi.NORM
    assert n_o[0, 0] > 1000.0
    np.testing.assert_allclose(n_d[0], n_o[0], rtol=1e-12)
