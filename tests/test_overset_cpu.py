"""Overset meshes on the CPU: the oracle's do_interpolations against the
reference (bit for bit over a 10-step SST + WENO5 run of a near-body C-grid
in a Cartesian background), and libot2d's host hole cutting (connect2d)
against the reference's lists, bit for bit, at the two-mesh airfoil sizes
385x97 + 401x401.  The hole cutting is pure host code: no GPU is needed."""
import os

import numpy as np
import pytest

from tests import helpers as H
from tests import overset_helpers as O
from turns2d_urans_git_amd import abi

LIB = os.path.join(H.ROOT, "turns2d_urans_git_amd", "libot2d.so")
FIXTURES = {"overset_97x25_61x61": ((97, 25), (61, 61)), "ihc_385x97_401x401": ((385, 97), (401, 401))}


def test_oracle_overset_trajectory_bitexact():
    z, deck = O.load("overset_97x25_61x61")
    ms = O.meshes(z)
    case = O.make(H.oracle_lib(), "or_", ms, deck)
    O.install_reference_connectivity(case, z)
    n = int(z["nsteps"])
    rc, hist, done = case.iterate(n)
    assert rc == 0 and done == n
    np.testing.assert_array_equal(hist, O.reference_norms(z, n))
    for im in (1, 2):
        np.testing.assert_array_equal(case.get_state(im - 1).ravel(order="F"), z[f"qfinal{im}"])
        np.testing.assert_array_equal(case.get_field(im - 1, abi.F_TURMU).ravel(order="F"), z[f"turmufinal{im}"])


def test_oracle_interpolation_moves_fringes_only():
    """do_interpolations rewrites the fringe cells and nothing else."""
    z, deck = O.load("overset_97x25_61x61")
    case = O.make(H.oracle_lib(), "or_", O.meshes(z), deck)
    O.install_reference_connectivity(case, z)
    rng = np.random.default_rng(5)
    before = []
    for im in range(2):
        q = case.get_state(im)
        q[..., :4] *= 1.0 + 0.01 * rng.standard_normal(q[..., :4].shape)
        case.set_state(im, q)
        before.append(q)
    case.interpolate()
    for im in range(2):
        after = case.get_state(im)
        jd = after.shape[0]
        fr = z[f"imesh{im + 1}"]
        nf = fr.size // 2
        mask = np.zeros(after.shape[:2], dtype=bool)
        mask[fr[:nf] - 1, fr[nf:] - 1] = True
        changed = np.any(after != before[im], axis=(2, 3))
        assert not np.any(changed & ~mask)
        assert np.all(changed[mask])
        assert jd > 0


def _ihc_inputs(z, deck, sizes):
    """Cell centres and areas of the oracle's initialisation (bit-identical to
    the reference's, tests/test_oracle_golden.py) as connect2d receives them."""
    ms = O.meshes(z, *sizes)
    case = O.make(H.oracle_lib(), "or_", ms, deck)
    ins = []
    for im, (xv, yv, _, ihc) in enumerate(ms):
        x = case.get_field(im, abi.F_X)[:, :, 0]
        y = case.get_field(im, abi.F_Y)[:, :, 0]
        qj = case.get_state(im)[:, :, 0, -1]
        ins.append(dict(xv=xv, yv=yv, xc=x, yc=y, vol=1.0 / qj, ihc=ihc))
    return ins


@pytest.mark.skipif(not os.path.exists(LIB), reason="libot2d.so not built")
@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_hole_cutting_matches_reference(name):
    import turns2d_urans_git_amd as t2d
    z, deck = O.load(name)
    lists = abi.ihc_connect(t2d.load_library(), "ot2d_", _ihc_inputs(z, deck, FIXTURES[name]),
                            errfn=t2d.last_error)
    for im in (1, 2):
        assert lists.counts(im - 1) == tuple(int(v) for v in z[f"counts{im}"])
        got = lists.get(im - 1)
        for key in ("iblank", "imesh", "ibc", "idonor", "frac"):
            np.testing.assert_array_equal(got[key], z[f"{key}{im}"], err_msg=f"mesh {im} {key}")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libot2d.so not built")
def test_fraction_roundtrip_and_hole():
    """Independent of the reference: the 9-point quadratic map of each donor
    sends its fractions back to the receiver's centre (Newton converged,
    fractions not clamped), every fringe points at an existing donor, and the
    blanked background cells lie inside the airfoil."""
    import turns2d_urans_git_amd as t2d
    z, deck = O.load("ihc_385x97_401x401")
    ins = _ihc_inputs(z, deck, FIXTURES["ihc_385x97_401x401"])
    lists = abi.ihc_connect(t2d.load_library(), "ot2d_", ins, errfn=t2d.last_error)
    got = [lists.get(i) for i in range(2)]
    nd = [g["idonor"].size // 2 for g in got]
    base = [0, nd[0]]
    checked = 0
    for n in range(2):
        g = got[n]
        nf = g["ibc"].size
        assert nf > 0
        for f in range(nf):
            gidx = int(g["ibc"][f]) - 1
            n1 = 0 if gidx < nd[0] else 1
            d = gidx - base[n1]
            assert 0 <= d < nd[n1]
            dj, dk = int(got[n1]["idonor"][d]), int(got[n1]["idonor"][nd[n1] + d])
            f1, f2 = got[n1]["frac"][d], got[n1]["frac"][nd[n1] + d]
            if abs(f1) >= 0.99 or abs(f2) >= 0.99:
                continue
            rj, rk = int(g["imesh"][f]), int(g["imesh"][nf + f])
            xc, yc = ins[n1]["xc"], ins[n1]["yc"]
            lj = {-1: 0.5 * f1 * (f1 - 1), 0: (1 - f1) * (1 + f1), 1: 0.5 * f1 * (f1 + 1)}
            lk = {-1: 0.5 * f2 * (f2 - 1), 0: (1 - f2) * (1 + f2), 1: 0.5 * f2 * (f2 + 1)}
            px = sum(lj[a] * lk[b] * xc[dj - 1 + a, dk - 1 + b] for a in (-1, 0, 1) for b in (-1, 0, 1))
            py = sum(lj[a] * lk[b] * yc[dj - 1 + a, dk - 1 + b] for a in (-1, 0, 1) for b in (-1, 0, 1))
            tx, ty = ins[n]["xc"][rj - 1, rk - 1], ins[n]["yc"][rj - 1, rk - 1]
            h = abs(xc[dj, dk - 1] - xc[dj - 2, dk - 1]) + abs(yc[dj - 1, dk] - yc[dj - 1, dk - 2])
            assert np.hypot(px - tx, py - ty) <= 1e-9 * max(h, 1.0), (n, f)
            checked += 1
    assert checked > 1000
    # blanked background cells: inside the airfoil (x in (0,1), |y| below the
    # 12 % thick section's half thickness bound)
    ib = got[1]["iblank"].reshape(402, 402, order="F")
    xb, yb = ins[1]["xc"], ins[1]["yc"]
    hole = np.argwhere(ib[1:-1, 1:-1] == 0) + 1
    assert len(hole) > 0
    for j, k in hole:
        assert 0.0 < xb[j, k] < 1.0 and abs(yb[j, k]) < 0.1
