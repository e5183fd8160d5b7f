"""The oracle (C restatement) against the reference's own outputs.

The fixtures under tests/golden/ were produced by tests/golden/make_golden.py
from the reference Fortran built from its sources (oracle/_ref).  The
restatement follows the reference operand for operand, so the comparison is
bit-exact: stage arrays of one iteration, the 40-step residual history, the
forces and the final state.
"""
import ctypes as C

import numpy as np
import pytest

from tests import helpers as H
from turns2d_urans_git_amd import abi

CASES = ["euler_33x17", "sa_41x21", "weno_sa_41x21", "sst_41x21"]


def _turb(case):
    ol = H.oracle_lib()
    ol.or_turbulence.argtypes = [C.c_void_p]
    ol.or_turbulence.restype = C.c_int32
    assert ol.or_turbulence(case) == 0


@pytest.mark.parametrize("name", CASES)
def test_initialisation_bitexact(name):
    z, case = H.golden_case(H.oracle_lib(), "or_", name)
    jd, kd, nq, nv, _, _ = case.dims[0]
    for f, key in [(abi.F_X, "x"), (abi.F_Y, "y"), (abi.F_XX, "xx"), (abi.F_XY, "xy"),
                   (abi.F_YX, "yx"), (abi.F_YY, "yy")]:
        np.testing.assert_array_equal(case.get_field(0, f).ravel(order="F"), z[key])
    np.testing.assert_array_equal(case.get_state(0).ravel(order="F"), z["qinit"])


@pytest.mark.parametrize("name", CASES)
def test_stages_bitexact(name):
    z, case = H.golden_case(H.oracle_lib(), "or_", name)
    lamin = name.startswith("euler")
    out = H.replay_stages(case, z, None if lamin else _turb)
    for st, a in out.items():
        np.testing.assert_array_equal(a.ravel(order="F"), z[st], err_msg=st)


@pytest.mark.parametrize("name", CASES)
def test_trajectory_bitexact(name):
    z, case = H.golden_case(H.oracle_lib(), "or_", name)
    n = int(z["nsteps"])
    norms = z["norms"]
    forces = z["forces"]
    for it in range(n):
        rc, nrm, am = case.step()
        assert rc == 0
        # full-precision norms of the reference (rsum, resrho, resmax, argmax)
        np.testing.assert_array_equal(nrm[0], norms[it, 3:6])
        np.testing.assert_array_equal(am[0], norms[it, 6:9].astype(int))
        f = case.forces()
        np.testing.assert_array_equal([f["cl"], f["cd"], f["cm"], f["cpower"], f["cfx"]], forces[it, 1:6])
    np.testing.assert_array_equal(case.get_state(0).ravel(order="F"), z["qfinal"])
    np.testing.assert_array_equal(case.get_field(0, abi.F_TURMU).ravel(order="F"), z["turmufinal"])


TS_CASES = ["ts3_euler_41x21", "ts4_sa_41x21"]


@pytest.mark.parametrize("name", TS_CASES)
def test_time_spectral_trajectory_bitexact(name):
    """Time-spectral coupling (src/sources.f:2-27) with pitching instances
    (src/setup.f:348-393, 727-786): every instance's norms and the final
    state of all instances, bit for bit."""
    z, case = H.golden_case(H.oracle_lib(), "or_", name)
    n, ns = int(z["nsteps"]), int(z["deck_nspec"])
    hist = np.array([case.step()[1].copy() for _ in range(n)]).reshape(n * ns, 3)
    np.testing.assert_array_equal(hist, z["norms"][:, 3:6])
    np.testing.assert_array_equal(case.get_state(0).ravel(order="F"), z["qfinal"])
