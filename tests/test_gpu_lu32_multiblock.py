"""FP32 LU-SGS sweeps whose rows span several workgroups.

The FP32 sweep (k_lu_sweep32) cuts the rows of a mesh into W = ceil(nk/32)
workgroups that pass each column's edge row to the next one through the
tagged inter-workgroup mail.  These tests cover W >= 2, with full and ragged
last blocks, against the all-FP64 oracle (ilu2d, src/lhs.f:470-814, in the
reference's operation order):

* one ilu2d application on an evolved SA state, dq cell by cell within
  1e-5 of the oracle's FP64 dq (relative to |dq| + max|dq|: the FP32
  operator is rounded once from FP64 and the sweep accumulates in FP32);
  a wrong or stale edge value would be off by O(1) along a whole row;
* 25-step trajectories: residual history within 1e-3 and state within 1e-5
  (max-norm relative) of the oracle, the bounds of the single-workgroup
  FP32 tests;
* an FP32 batch of 8 cases of the 64-case mix with kmax spread over 66-193
  (W = 3 to 6, uneven jmax), every case against the oracle;
* all 64 cases of the mix for 20 steps: every case in the batch equals the
  same case alone, bit for bit;
* three FP32 batches of the mix overlapping on three streams for 2000 steps:
  every case equals its run on one stream with nothing beside it, bit for
  bit, checked every 250 steps (the test stops at the first mismatch).
"""
import numpy as np
import pytest

import bench
from tests import helpers as H
from turns2d_urans_git_amd import abi, grid

pytestmark = pytest.mark.gpu

SA_DECK = dict(fsmach=0.7, alfa=2.0, rey=9.0e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3,
               ilim=1, ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1,
               timeac=1.0)

ROWS = 32   # rows of one sweep workgroup

# (jmax, kmax): nk = kmax-1 rows -> W blocks, rows in the last block
GRIDS = [
    (129, 67),    # nk 66: W 3, last block 2 rows
    (129, 66),    # nk 65: W 3, last block 1 row
    (257, 131),   # nk 130: W 5, last block 2 rows
    (257, 193),   # nk 192: W 6, full blocks
]


class _Stages(dict):
    """Input of helpers.replay_stages in the form of a fixture file."""

    @property
    def files(self):
        return list(self.keys())


def _evolved(g, deck, n=10):
    """The oracle after n steps from the free stream, and its state as the
    input of one replayed step."""
    o = H.oracle_case(g, deck)
    rc, _, _ = o.iterate(n)
    assert rc == 0
    return o, _Stages(qin=o.get_state(0).ravel(order="F"), turmuin=o.get_field(0, abi.F_TURMU).ravel(order="F"),
                      dumpat=n)


@pytest.mark.parametrize("jm,km", GRIDS)
def test_lu_stage_vs_fp64_oracle(jm, km):
    g = H.cgrid(jm, km, ds_wall=1e-5)
    print(f"[lu32 stage] {jm}x{km} W={(km - 1 + ROWS - 1) // ROWS}", flush=True)
    o, z = _evolved(g, SA_DECK)
    d = H.gpu_case(g, dict(SA_DECK, lu_fp32=1))
    ref = H.replay_stages(o, z)
    got = H.replay_stages(d, z)
    # the stages before ilu2d are FP64 in both
    H.assert_cellwise(got["rbc"], ref["rbc"], 1e-12, "residual after rbc")
    dq, dq_ref = got["dq"], ref["dq"]
    assert np.abs(dq_ref).max() > 0
    H.assert_cellwise(dq, dq_ref, 1e-5, f"dq after the FP32 sweeps ({jm}x{km})")
    # the rows next to every block edge carry values, not zeros
    jd, kd = dq.shape[:2]
    for w in range(1, (km - 1 + ROWS - 1) // ROWS):
        k = 1 + ROWS * w   # first row of block w (0-based k of the halo layout)
        assert np.abs(dq[1:jd - 1, k, 0, :]).max() > 0, (w, k)


@pytest.mark.parametrize("jm,km", GRIDS)
def test_trajectory_vs_fp64_oracle(jm, km):
    g = H.cgrid(jm, km, ds_wall=1e-5)
    print(f"[lu32 trajectory] {jm}x{km}", flush=True)
    o = H.oracle_case(g, SA_DECK)
    d = H.gpu_case(g, dict(SA_DECK, lu_fp32=1))
    rc_o, h_o, n_o = o.iterate(25)
    rc_d, h_d, n_d = d.iterate(25)
    assert rc_o == rc_d == 0 and n_o == n_d == 25
    np.testing.assert_allclose(h_d, h_o, rtol=1e-3)
    assert H.rel_err(d.get_state(0), o.get_state(0)) < 1e-5


def test_euler_multiblock_vs_fp64_oracle():
    """Euler (the update fused into the sweep's finish kernel), W = 3."""
    g = H.cgrid(129, 97, ds_wall=1e-3)
    o = H.oracle_case(g, H.EULER)
    d = H.gpu_case(g, dict(H.EULER, lu_fp32=1))
    rc_o, h_o, _ = o.iterate(25)
    rc_d, h_d, _ = d.iterate(25)
    assert rc_o == rc_d == 0
    np.testing.assert_allclose(h_d, h_o, rtol=1e-3)
    assert H.rel_err(d.get_state(0), o.get_state(0)) < 1e-5


def _mix(indices=None, lu_fp32=1):
    import turns2d_urans_git_amd as t2d
    cases = bench.batch_cases()
    idx = list(range(len(cases))) if indices is None else list(indices)
    grids, params, meshes, decks = [], [], [], []
    for i in idx:
        c = cases[i]
        g = grid.naca4_cgrid(c["jmax"], c["kmax"], m=c["m"], p=c["p"], t=c["t"], ds_wall=1e-6,
                             wake_aspect=1000.0, te_cluster=False)
        deck = dict(SA_DECK, fsmach=c["fsmach"], alfa=c["alfa"], rey=c["rey"])
        grids.append(g)
        decks.append(deck)
        params.append(t2d.default_params(**deck, lu_fp32=lu_fp32))
        meshes.append(abi.mesh_desc(g.jmax, g.kmax, grid.cgrid_bc(g.jtail1)))
    return idx, grids, decks, params, meshes


def _spread_by_kmax(n):
    cases = bench.batch_cases()
    order = sorted(range(len(cases)), key=lambda i: cases[i]["kmax"])
    return [order[int(round(x))] for x in np.linspace(0, len(order) - 1, n)]


def test_fp32_batch_cases_vs_fp64_oracle():
    import turns2d_urans_git_amd as t2d
    idx, grids, decks, params, meshes = _mix(_spread_by_kmax(8))
    km = sorted(g.kmax for g in grids)
    assert km[0] <= 80 and km[-1] >= 180, km
    print(f"[lu32 batch] cases {idx} kmax {km}", flush=True)
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    n = 20
    rc, hist, done = b.iterate(n)
    assert rc == 0 and done == n
    st, ist = b.batch_status()
    assert list(st) == [0] * len(idx) and list(ist) == [n] * len(idx)
    for i, g in enumerate(grids):
        o = H.oracle_case(g, decks[i])
        rc_o, h_o, n_o = o.iterate(n)
        assert rc_o == 0 and n_o == n
        np.testing.assert_allclose(hist[:, i], h_o[:, 0], rtol=1e-3, err_msg=f"case {idx[i]}")
        e = H.rel_err(b.get_state(i), o.get_state(0))
        assert e < 1e-5, (idx[i], g.jmax, g.kmax, e)


def test_full_mix_batch_equals_alone():
    """All 64 cases for 20 steps, FP32 LU-SGS: batch == alone, bit for bit."""
    import turns2d_urans_git_amd as t2d
    idx, grids, decks, params, meshes = _mix()
    b = t2d.Batch(params, meshes)
    for i, g in enumerate(grids):
        b.set_grid(i, g.xv, g.yv)
    n = 20
    rc, hist, done = b.iterate(n)
    assert rc == 0 and done == n
    for i, g in enumerate(grids):
        one = t2d.Case(params[i], [meshes[i]])
        one.set_grid(0, g.xv, g.yv)
        rc1, h1, _ = one.iterate(n)
        assert rc1 == 0
        np.testing.assert_array_equal(hist[:, i], h1[:, 0], err_msg=f"case {i}")
        np.testing.assert_array_equal(b.get_state(i), one.get_state(0), err_msg=f"case {i}")
        one.close()
    b.close()


@pytest.mark.parametrize("ordered", [True, False])
def test_overlapped_fp32_batches_equal_one_stream(ordered):
    """Three FP32 batches of the mix (W = 3 to 6 cases) enqueued on three
    streams at once, against the same three batches each run with nothing
    beside it, bit for bit, checked every `chunk` steps: with the library's
    order of FP32 work across streams (the default) and with it lifted
    (ot2d_set_fp32_order(0)), where the three batches' sweeps run at the
    same time."""
    import turns2d_urans_git_amd as t2d
    lib = t2d.load_library()
    idx, grids, decks, params, meshes = _mix()
    cells = np.array([(g.jmax - 1) * (g.kmax - 1) for g in grids], dtype=np.float64)
    groups = bench.batch_groups(cells, 3)

    def make(grp):
        b = t2d.Batch([params[i] for i in grp], [meshes[i] for i in grp])
        for n, i in enumerate(grp):
            b.set_grid(n, grids[i].xv, grids[i].yv)
        return b

    over = [make(g) for g in groups]
    alone = [make(g) for g in groups]
    chunk, nchunk = 250, 8
    abi.set_fp32_order(lib, "ot2d_", ordered)
    try:
        for c in range(nchunk):
            for b in over:
                b.iterate_async(chunk)
            for b in over:
                b.sync()
            for b in alone:          # one at a time: nothing runs beside it
                b.iterate_async(chunk)
                b.sync()
            for gi, grp in enumerate(groups):
                st_o, is_o = over[gi].batch_status()
                st_a, is_a = alone[gi].batch_status()
                np.testing.assert_array_equal(st_o, st_a, err_msg=f"status, group {gi}, chunk {c}")
                np.testing.assert_array_equal(is_o, is_a, err_msg=f"steps, group {gi}, chunk {c}")
                for n, i in enumerate(grp):
                    qo, qa = over[gi].get_state(n), alone[gi].get_state(n)
                    assert np.array_equal(qo, qa, equal_nan=True), \
                        f"case {i} (group {gi}) differs after {(c + 1) * chunk} steps (ordered={ordered})"
            print(f"[lu32 overlap ordered={ordered}] {(c + 1) * chunk} steps: all {len(idx)} cases equal",
                  flush=True)
    finally:
        abi.set_fp32_order(lib, "ot2d_", True)
    for b in over + alone:
        b.close()
