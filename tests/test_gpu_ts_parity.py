"""GPU parity of the time-spectral iteration (pitching airfoil, N instances
coupled by the spectral time derivative, src/sources.f:2-27) against the
reference's own trajectories (tests/golden/ts*_41x21.npz).

Tolerance: per-instance residual norms of every step to 1e-10 relative and
the final state of all instances per cell to 1e-10 (FP64 throughout)."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["ts3_euler_41x21", "ts4_sa_41x21"])
def test_time_spectral_vs_reference(name):
    import turns2d_urans_git_amd as t2d
    z, case = H.golden_case(t2d.load_library(), "ot2d_", name, errfn=t2d.last_error)
    n, ns = int(z["nsteps"]), int(z["deck_nspec"])
    rc, hist, done = case.iterate(n)
    assert rc == 0 and done == n
    np.testing.assert_allclose(hist.reshape(n * ns), z["norms"][:, 3], rtol=1e-10)
    H.assert_cellwise(case.get_state(0).ravel(order="F"), z["qfinal"], 1e-10, "q (all instances)")


@pytest.mark.parametrize("name", ["ts3_euler_41x21", "ts4_sa_41x21"])
def test_time_spectral_step_norms(name):
    """ot2d_step reports rsum, resrho, resmax of every instance."""
    import turns2d_urans_git_amd as t2d
    z, case = H.golden_case(t2d.load_library(), "ot2d_", name, errfn=t2d.last_error)
    ns = int(z["deck_nspec"])
    for it in range(4):
        rc, nrm, _ = case.step()
        assert rc == 0
        np.testing.assert_allclose(nrm, z["norms"][it * ns:(it + 1) * ns, 3:6], rtol=1e-10)


def test_time_spectral_fp32_lusgs():
    import turns2d_urans_git_amd as t2d
    z, case = H.golden_case(t2d.load_library(), "ot2d_", "ts4_sa_41x21", errfn=t2d.last_error, lu_fp32=1)
    n, ns = int(z["nsteps"]), int(z["deck_nspec"])
    rc, hist, done = case.iterate(n)
    assert rc == 0 and done == n
    np.testing.assert_allclose(hist.reshape(n * ns), z["norms"][:, 3], rtol=1e-3)
    assert H.rel_err(case.get_state(0).ravel(order="F"), z["qfinal"]) < 1e-5
