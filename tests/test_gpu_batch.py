"""Batches (ot2d_batch_create): independent cases of uneven size and
different flow conditions advanced by the same launches.  Every case's
result must be the same in the batch as alone (bit for bit: the kernels
compute each instance from its own data in a fixed order), including after
another case of the batch stops or is retired."""
import numpy as np
import pytest

from tests import helpers as H
from turns2d_urans_git_amd import abi, grid

pytestmark = pytest.mark.gpu

SA = dict(fsmach=0.5, alfa=2.0, rey=1.0e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3, ilim=1,
          ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0)
CASES = [
    # (jmax, kmax, camber m, p, t, ds_wall, deck overrides)
    (65, 25, 0.0, 0.0, 0.12, 1e-4, dict(fsmach=0.5, alfa=2.0)),
    (97, 33, 0.02, 0.4, 0.10, 1e-4, dict(fsmach=0.7, alfa=-1.5, rey=3.0e6)),
    (81, 73, 0.04, 0.3, 0.15, 2e-4, dict(fsmach=0.3, alfa=5.0, rey=5.0e5)),
]


def _inputs(lu_fp32):
    import turns2d_urans_git_amd as t2d
    grids, params, meshes = [], [], []
    for jm, km, m, p, t, ds, over in CASES:
        g = grid.naca4_cgrid(jm, km, m=m, p=p, t=t, ds_wall=ds)
        grids.append(g)
        params.append(t2d.default_params(**dict(SA, **over), lu_fp32=lu_fp32))
        meshes.append(abi.mesh_desc(jm, km, grid.cgrid_bc(g.jtail1)))
    return grids, params, meshes


@pytest.mark.parametrize("lu_fp32", [0, 1])
def test_batch_equals_alone(lu_fp32):
    import turns2d_urans_git_amd as t2d
    grids, params, meshes = _inputs(lu_fp32)
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    n = 25
    rc, hist, done = b.iterate(n)
    assert rc == 0 and done == n
    st, ist = b.batch_status()
    assert list(st) == [0, 0, 0] and list(ist) == [n, n, n]
    for i, g in enumerate(grids):
        solo = t2d.Case(params[i], [meshes[i]])
        solo.set_grid(0, g.xv, g.yv)
        rc_s, h_s, _ = solo.iterate(n)
        assert rc_s == 0
        np.testing.assert_array_equal(hist[:, i], h_s[:, 0])
        np.testing.assert_array_equal(b.get_state(i), solo.get_state(0))
        np.testing.assert_array_equal(b.get_field(i, abi.F_TURMU), solo.get_field(0, abi.F_TURMU))
        fb, fs = b.forces(i), solo.forces(0)
        assert fb == fs


def test_batch_first_case_against_oracle():
    import turns2d_urans_git_amd as t2d
    grids, params, meshes = _inputs(0)
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    rc, hist, _ = b.iterate(20)
    jm, km, m, p, t, ds, over = CASES[1]
    o = H.oracle_case(grids[1], dict(SA, **over))
    _, h_o, _ = o.iterate(20)
    np.testing.assert_allclose(hist[:, 1], h_o[:, 0], rtol=1e-10)
    H.assert_cellwise(b.get_state(1), o.get_state(0), 1e-10, "q of case 1 in the batch")


def test_retired_case_is_frozen_and_others_unaffected():
    import turns2d_urans_git_amd as t2d
    grids, params, meshes = _inputs(1)
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    b.iterate(10)
    q1 = b.get_state(1).copy()
    b.batch_retire(1)
    rc, hist, done = b.iterate(15)
    assert rc == 0 and done == 15
    st, ist = b.batch_status()
    assert list(st) == [0, abi.CONVERGED, 0] and list(ist) == [25, 10, 25]
    np.testing.assert_array_equal(b.get_state(1), q1)
    assert np.all(hist[:, 1] == 0.0)          # no norms recorded for a retired case
    solo = t2d.Case(params[2], [meshes[2]])
    solo.set_grid(0, grids[2].xv, grids[2].yv)
    solo.iterate(25)
    np.testing.assert_array_equal(b.get_state(2), solo.get_state(0))
    # setting the grid again restarts the retired case
    b.set_grid(1, grids[1].xv, grids[1].yv)
    st, ist = b.batch_status()
    assert st[1] == 0 and ist[1] == 0


def test_many_line_batch_equals_alone():
    """A batch with more than 4096 grid lines runs the SA line solves one lane
    per line (tiled through LDS along j, row-contiguous along k); each case
    must still match its solo run (one wave per line) bit for bit.  The grid
    sizes leave partial last tiles along j (j lines of 134, 100, 76 cells),
    a line shorter than one tile (24 cells), and partial row blocks (69 and
    65 rows: a second block of 5 and of 1 row)."""
    import turns2d_urans_git_amd as t2d
    rng = np.random.default_rng(7)
    sizes = [(129, 49), (135, 70), (101, 33), (77, 66), (25, 17)]
    grids, params, meshes = [], [], []
    for n in range(32):
        jm, km = sizes[n % len(sizes)]
        g = grid.naca4_cgrid(jm, km, m=float(rng.uniform(0, 0.05)), p=0.4, t=float(rng.uniform(0.09, 0.15)),
                             ds_wall=1e-4)
        grids.append(g)
        params.append(t2d.default_params(**dict(SA, fsmach=float(rng.uniform(0.3, 0.7)),
                                                alfa=float(rng.uniform(-2, 5))), lu_fp32=1))
        meshes.append(abi.mesh_desc(jm, km, grid.cgrid_bc(g.jtail1)))
    assert sum((jm - 1) + (km - 1) for jm, km in (sizes[n % len(sizes)] for n in range(32))) > 4096
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    rc, hist, done = b.iterate(15)
    assert rc == 0 and done == 15
    for i in (0, 1, 2, 3, 4, 31):
        solo = t2d.Case(params[i], [meshes[i]])
        solo.set_grid(0, grids[i].xv, grids[i].yv)
        _, h_s, _ = solo.iterate(15)
        np.testing.assert_array_equal(hist[:, i], h_s[:, 0])
        np.testing.assert_array_equal(b.get_state(i), solo.get_state(0))
        np.testing.assert_array_equal(b.get_field(i, abi.F_TURMU), solo.get_field(0, abi.F_TURMU))


def test_nonfinite_residual_stops_with_norm():
    """With the convergence test on, a case whose residual is not a finite
    number stops with NORM, not CONVERGED; the other cases run on."""
    import turns2d_urans_git_amd as t2d
    grids, params, meshes = _inputs(0)
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    b.set_convergence(1e-10)
    q = b.get_state(1)
    q[5:9, 5:9, :, :4] = np.nan
    b.set_state(1, q)
    b.iterate_async(3)
    st, ist = b.batch_status()
    assert st[1] == abi.NORM, st
    assert st[0] == abi.OK and st[2] == abi.OK, st
    assert ist[0] == 3 and ist[2] == 3


@pytest.mark.parametrize("lu_fp32", [0, 1])
def test_compacted_batch_runs_only_running_cases(lu_fp32):
    """ot2d_batch_compact: after a case is retired, the batch's steps launch
    over the running cases' records and tiles only.  The running cases'
    histories and states equal those of an uncompacted twin bit for bit, the
    retired case stays frozen, and setting its grid again returns the batch
    to every case (the case runs again, as in the twin)."""
    import turns2d_urans_git_amd as t2d
    grids, params, meshes = _inputs(lu_fp32)

    def make():
        b = t2d.Batch(params, meshes)
        for i, g in enumerate(grids):
            b.set_grid(i, g.xv, g.yv)
        b.iterate(8)
        b.batch_retire(0)
        return b

    comp, twin = make(), make()
    q0 = comp.get_state(0).copy()
    assert comp.batch_compact() == 2
    assert comp.batch_compact() == 2          # unchanged: nothing to redo
    r1, h1, d1 = comp.iterate(12)
    r2, h2, d2 = twin.iterate(12)
    assert r1 == r2 == 0 and d1 == d2 == 12
    np.testing.assert_array_equal(h1, h2)
    assert np.all(h1[:, 0] == 0.0)
    np.testing.assert_array_equal(comp.get_state(0), q0)
    for i in (1, 2):
        np.testing.assert_array_equal(comp.get_state(i), twin.get_state(i), err_msg=f"case {i}")
    assert list(comp.batch_status()[1]) == list(twin.batch_status()[1]) == [8, 20, 20]
    # async steps on the compacted records as well
    comp.iterate_async(5)
    twin.iterate_async(5)
    comp.sync()
    twin.sync()
    for i in (1, 2):
        np.testing.assert_array_equal(comp.get_state(i), twin.get_state(i), err_msg=f"case {i}, async")
    # the retired case runs again after its grid is set: every case launches
    for b in (comp, twin):
        b.set_grid(0, grids[0].xv, grids[0].yv)
        rc, _, _ = b.iterate(6)
        assert rc == 0
    for i in range(3):
        np.testing.assert_array_equal(comp.get_state(i), twin.get_state(i), err_msg=f"case {i}, restarted")
    # every case running: compaction keeps the full records
    assert comp.batch_compact() == 3
    comp.close()
    twin.close()
