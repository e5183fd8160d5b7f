"""GPU parity of the overset path: connectivity found from the device's own
grid (ot2d_connect) against the reference's connect2d lists, and the
SST + WENO5 iteration with do_interpolations after every step against the
reference's 10-step run (1e-10) and the oracle (FP32 LU-SGS, 1e-3)."""
import numpy as np
import pytest

from tests import helpers as H
from tests import overset_helpers as O
from turns2d_urans_git_amd import abi

pytestmark = pytest.mark.gpu


def _gpu(ms, deck):
    import turns2d_urans_git_amd as t2d
    return O.make(t2d.load_library(), "ot2d_", ms, deck, errfn=t2d.last_error)


@pytest.mark.parametrize("name,sizes", [("overset_97x25_61x61", (None, None)),
                                        ("ihc_385x97_401x401", ((385, 97), (401, 401)))])
def test_connect_on_device_matches_reference(name, sizes):
    z, deck = O.load(name)
    ms = O.meshes(z, *sizes)
    case = _gpu(ms, deck)
    nfr = case.connect([m[3] for m in ms])
    assert nfr == int(z["counts1"][0]) + int(z["counts2"][0])
    for im in (1, 2):
        ib = case.get_iblank(im - 1).ravel(order="F")
        np.testing.assert_array_equal(ib, z[f"iblank{im}"], err_msg=f"mesh {im}")


def test_overset_trajectory_vs_reference():
    z, deck = O.load("overset_97x25_61x61")
    ms = O.meshes(z)
    case = _gpu(ms, deck)
    case.connect([m[3] for m in ms])
    n = int(z["nsteps"])
    rc, hist, done = case.iterate(n)
    assert rc == 0 and done == n
    np.testing.assert_allclose(hist, O.reference_norms(z, n), rtol=1e-10)
    for im in (1, 2):
        H.assert_cellwise(case.get_state(im - 1).ravel(order="F"), z[f"qfinal{im}"], 1e-10, f"q mesh {im}")
        H.assert_cellwise(case.get_field(im - 1, abi.F_TURMU).ravel(order="F"), z[f"turmufinal{im}"], 1e-10,
                          f"turmu mesh {im}")


def test_overset_fp32_lusgs_vs_oracle():
    z, deck = O.load("overset_97x25_61x61")
    ms = O.meshes(z)
    o = O.make(H.oracle_lib(), "or_", ms, deck)
    O.install_reference_connectivity(o, z)
    d = _gpu(ms, dict(deck, lu_fp32=1))
    d.connect([m[3] for m in ms])
    rc_o, h_o, _ = o.iterate(10)
    rc_d, h_d, _ = d.iterate(10)
    assert rc_o == rc_d == 0
    np.testing.assert_allclose(h_d, h_o, rtol=1e-3)
    for im in range(2):
        assert H.rel_err(d.get_state(im), o.get_state(im)) < 1e-5


def test_interpolate_entry_point_matches_oracle():
    """ot2d_interpolate alone on a perturbed state equals the oracle's
    do_interpolations on the same state."""
    z, deck = O.load("overset_97x25_61x61")
    ms = O.meshes(z)
    o = O.make(H.oracle_lib(), "or_", ms, deck)
    O.install_reference_connectivity(o, z)
    d = _gpu(ms, deck)
    d.connect([m[3] for m in ms])
    rng = np.random.default_rng(11)
    for im in range(2):
        q = o.get_state(im)
        q[..., :4] *= 1.0 + 0.01 * rng.standard_normal(q[..., :4].shape)
        q[..., 4:6] *= 1.0 + 0.1 * rng.random(q[..., 4:6].shape)
        o.set_state(im, q)
        d.set_state(im, q)
    o.interpolate()
    d.interpolate()
    for im in range(2):
        H.assert_cellwise(d.get_state(im), o.get_state(im), 1e-14, f"mesh {im}")
