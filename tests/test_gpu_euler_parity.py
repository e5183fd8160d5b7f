"""GPU parity of the Euler iteration (plain FP64 path) against the oracle.

Every stage entry point of libot2d is compared with the oracle's restatement
of the same reference subroutine on the same state.  Tolerance: per cell
|gpu - oracle| <= 1e-12 * (|oracle| + max|oracle|) (FP64; the kernels keep
the reference's operand order, so most arrays agree bit for bit).  Norms
(reordered reductions) are checked to 1e-12 relative.
"""
import numpy as np
import pytest

from tests import helpers as H
from turns2d_urans_git_amd import abi

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def pair():
    g = H.cgrid(65, 25, ds_wall=1e-3)
    o = H.oracle_case(g, H.EULER)
    d = H.gpu_case(g, H.EULER)
    # a developed state: 25 oracle steps from free stream (inside the wall ramp)
    for _ in range(25):
        o.step()
    return g, o, d


GEOM = [("x", abi.F_X), ("y", abi.F_Y), ("xx", abi.F_XX), ("xy", abi.F_XY),
        ("yx", abi.F_YX), ("yy", abi.F_YY)]


def test_geometry_and_freestream():
    g = H.cgrid(65, 25, ds_wall=1e-3)
    o = H.oracle_case(g, H.EULER)
    d = H.gpu_case(g, H.EULER)
    for name, f in GEOM:
        H.assert_cellwise(d.get_field(0, f), o.get_field(0, f), RTOL, name)
    H.assert_cellwise(d.get_state(0), o.get_state(0), RTOL, "q after initia")


def _sync(o, d):
    d.set_state(0, o.get_state(0))
    d.istep0 = o.istep0
    for f in (abi.F_S,):
        d.set_field(0, f, o.get_field(0, f))


def test_time_step(pair):
    g, o, d = pair
    _sync(o, d)
    o.time_step()
    d.time_step()
    H.assert_cellwise(d.get_field(0, abi.F_TSCALE), o.get_field(0, abi.F_TSCALE), RTOL, "tscale")


def test_bc(pair):
    g, o, d = pair
    _sync(o, d)
    o.bc()
    d.bc()
    H.assert_cellwise(d.get_state(0), o.get_state(0), RTOL, "q after bc")


def test_residual_rbc_lhs(pair):
    g, o, d = pair
    _sync(o, d)
    o.time_step(); d.time_step()
    z = np.zeros_like(o.get_field(0, abi.F_S))
    o.set_field(0, abi.F_S, z); d.set_field(0, abi.F_S, z)
    o.compute_residual(); d.compute_residual()
    H.assert_cellwise(d.get_state(0), o.get_state(0), RTOL, "q after bc")
    s_o = o.get_field(0, abi.F_S)
    assert np.abs(s_o).max() > 1e-8  # a non-trivial residual
    H.assert_cellwise(d.get_field(0, abi.F_S), s_o, RTOL, "residual")
    o.rbc(); d.rbc()
    H.assert_cellwise(d.get_field(0, abi.F_S), o.get_field(0, abi.F_S), RTOL, "rbc")
    o.lhs(); d.lhs()
    H.assert_cellwise(d.get_field(0, abi.F_S), o.get_field(0, abi.F_S), RTOL, "dq (ilu2d)")
    H.assert_cellwise(d.get_state(0), o.get_state(0), RTOL, "q after ilu2d round trip")


def test_step_and_norms(pair):
    g, o, d = pair
    _sync(o, d)
    rc_o, n_o, a_o = o.step()
    rc_d, n_d, a_d = d.step()
    assert rc_o == rc_d == 0
    H.assert_cellwise(d.get_state(0), o.get_state(0), RTOL, "q after step")
    np.testing.assert_allclose(n_d, n_o, rtol=1e-12)
    np.testing.assert_array_equal(a_d, a_o)
    assert d.istep0 == o.istep0


def test_iterate_trajectory():
    """60 steps from free stream (through the 30-step wall ramp)."""
    g = H.cgrid(65, 25, ds_wall=1e-3)
    o = H.oracle_case(g, H.EULER)
    d = H.gpu_case(g, H.EULER)
    rc_o, h_o, n_o = o.iterate(60)
    rc_d, h_d, n_d = d.iterate(60)
    assert rc_o == rc_d == 0 and n_o == n_d == 60
    np.testing.assert_allclose(h_d, h_o, rtol=1e-11)
    H.assert_cellwise(d.get_state(0), o.get_state(0), 1e-11, "q after 60 steps")
    fo, fd = o.forces(), d.forces()
    for k in ("cl", "cd", "cm"):
        assert abs(fd[k] - fo[k]) <= 1e-11 * max(abs(fo[k]), 1e-3), (k, fd[k], fo[k])


def test_fp32_lusgs_defect_correction():
    """FP32 LU-SGS (lu_fp32=1): the residual stays FP64, so the update
    direction differs only at FP32 rounding; after 60 steps the state stays
    within 1e-5 relative of the all-FP64 oracle, and the residual history
    tracks it to 1e-3."""
    g = H.cgrid(65, 25, ds_wall=1e-3)
    o = H.oracle_case(g, H.EULER)
    d = H.gpu_case(g, dict(H.EULER, lu_fp32=1))
    rc_o, h_o, _ = o.iterate(60)
    rc_d, h_d, _ = d.iterate(60)
    assert rc_o == rc_d == 0
    np.testing.assert_allclose(h_d, h_o, rtol=1e-3)
    assert H.rel_err(d.get_state(0), o.get_state(0)) < 1e-5


def test_history_after_buffer_growth():
    """A longer ot2d_iterate after a short one records every step's norm
    (the device history buffer is regrown between the calls)."""
    g = H.cgrid(65, 25, ds_wall=1e-3)
    o = H.oracle_case(g, H.EULER)
    d = H.gpu_case(g, H.EULER)
    _, h1o, _ = o.iterate(3)
    _, h2o, _ = o.iterate(40)
    rc1, h1, n1 = d.iterate(3)
    rc2, h2, n2 = d.iterate(40)
    assert rc1 == rc2 == 0 and n1 == 3 and n2 == 40
    assert np.all(h2[:, 0] > 0.0)
    np.testing.assert_allclose(h1, h1o, rtol=1e-10)
    np.testing.assert_allclose(h2, h2o, rtol=1e-10)
