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


def test_pitching_overset_iterate_equals_steps():
    """A pitching overset case (the near-body mesh turns every step, the
    connectivity is cut again after each step's motion): ot2d_iterate over
    four steps and four ot2d_step calls give the same states, grid and
    iblank bit for bit; the near-body grid turns and the background's does
    not; the iblank the steps leave is that of a fresh cut of the turned
    grids; the graph-only entry points refuse the case, and the launch search
    leaves its grid alone (no search)."""
    import turns2d_urans_git_amd as t2d
    from turns2d_urans_git_amd import grid
    gms = grid.airfoil_overset(near=(97, 25), back=(61, 61), ds_wall=1e-4)
    ms = [(m.xv, m.yv, m.bc, m.ihc) for m in gms]
    deck = dict(fsmach=0.3, alfa=2.0, rey=3.0e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3, ilim=1,
                ilhs=1, iconstant_cfl=0, dt=0.1, epse=0.01, iunst=1, motiontype=1, rf=1.0, angmax=4.0, ntac=1,
                itnmax=1, timeac=1.0)
    a, b = _gpu(ms, deck), _gpu(ms, deck)
    for c in (a, b):
        c.connect([m[3] for m in ms])
    xv0 = [a.get_field(im, abi.F_XV).copy() for im in (0, 1)]
    print("[pitch overset] iterate 4 / step x4", flush=True)
    rc, _, done = a.iterate(4)
    assert rc == 0 and done == 4
    for _ in range(4):
        b.step()
    for im in (0, 1):
        np.testing.assert_array_equal(a.get_state(im), b.get_state(im), err_msg=f"q of mesh {im}")
        np.testing.assert_array_equal(a.get_iblank(im), b.get_iblank(im), err_msg=f"iblank of mesh {im}")
        np.testing.assert_array_equal(a.get_field(im, abi.F_XV), b.get_field(im, abi.F_XV))
    assert not np.array_equal(a.get_field(0, abi.F_XV), xv0[0]), "the near-body mesh did not turn"
    np.testing.assert_array_equal(a.get_field(1, abi.F_XV), xv0[1], err_msg="the background moved")
    # the cut the steps left is the cut of the turned grid
    c = _gpu(ms, deck)
    for im in (0, 1):
        c.set_grid(im, a.get_field(im, abi.F_XV), a.get_field(im, abi.F_YV))
    c.connect([m[3] for m in ms])
    for im in (0, 1):
        np.testing.assert_array_equal(a.get_iblank(im), c.get_iblank(im), err_msg=f"cut of the turned mesh {im}")
    c.close()
    with pytest.raises(abi.CaseError):
        a.iterate_async(1)
    sw, sl, measured = a.autotune(steps=2)
    assert not measured
    np.testing.assert_array_equal(a.get_field(0, abi.F_XV), b.get_field(0, abi.F_XV))
    a.close()
    b.close()
