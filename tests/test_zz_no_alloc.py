"""ot2d_iterate_noalloc (include/ot2d.h): the same steps as ot2d_iterate
with the residual history staged in the library's fixed scratch array.
Called through the C-ABI and checked against ot2d_iterate on twin cases, bit
for bit (same kernels): two calls in a row, the second with a smaller case
and fewer steps, a continuation on the same case, a batch, and the size
limit."""
import ctypes as C

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

SA = dict(fsmach=0.7, alfa=2.0, rey=9.0e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3, ilim=1, ilhs=1,
          iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0)


def _fn():
    import turns2d_urans_git_amd as t2d
    f = t2d.load_library().ot2d_iterate_noalloc
    f.restype = C.c_int32
    f.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]
    return f


def _noalloc(case, n, per=1, hist=True):
    h = np.full(n * per, np.nan)
    done = C.c_int32(-1)
    rc = _fn()(case.h, n, h.ctypes.data if hist else None, C.byref(done))
    return rc, h.reshape(n, per)[: done.value], done.value


@pytest.mark.parametrize("lu_fp32", [0, 1])
def test_two_calls_match_iterate(lu_fp32):
    for jm, km, n in ((129, 67, 30), (65, 25, 10)):     # the second call: a smaller case, fewer steps
        g = H.cgrid(jm, km, ds_wall=1e-5)
        deck = dict(SA, lu_fp32=lu_fp32)
        ref, d = H.gpu_case(g, deck), H.gpu_case(g, deck)
        rc_r, h_r, n_r = ref.iterate(n)
        rc, h, done = _noalloc(d, n)
        assert rc == rc_r == 0 and done == n_r == n
        np.testing.assert_array_equal(h, h_r)
        np.testing.assert_array_equal(d.get_state(0), ref.get_state(0))
        # continuing on the same cases: noalloc after noalloc, and ot2d_iterate after noalloc
        rc_r, h_r, _ = ref.iterate(7)
        rc, h, done = _noalloc(d, 7)
        assert rc == rc_r == 0 and done == 7
        np.testing.assert_array_equal(h, h_r)
        rc_r, h_r, _ = ref.iterate(5)
        rc2, h2, _ = d.iterate(5)
        np.testing.assert_array_equal(h2, h_r)
        np.testing.assert_array_equal(d.get_state(0), ref.get_state(0))
        ref.close()
        d.close()


def test_batch_and_no_history():
    import turns2d_urans_git_amd as t2d
    from turns2d_urans_git_amd import abi, grid
    gs = [grid.naca4_cgrid(jm, km, ds_wall=1e-5) for jm, km in ((97, 73), (65, 33), (129, 49))]
    ps = [t2d.default_params(**dict(SA, alfa=a), lu_fp32=1) for a in (1.0, 2.0, 3.0)]
    mds = [abi.mesh_desc(g.jmax, g.kmax, grid.cgrid_bc(g.jtail1)) for g in gs]

    def mk():
        b = t2d.Batch(ps, mds)
        for i, g in enumerate(gs):
            b.set_grid(i, g.xv, g.yv)
        return b

    ref, d = mk(), mk()
    rc_r, h_r, n_r = ref.iterate(20)
    rc, h, done = _noalloc(d, 20, per=3)
    assert rc == rc_r == 0 and done == n_r == 20
    np.testing.assert_array_equal(h, h_r)
    # without a history
    ref.iterate(6)
    rc, _, done = _noalloc(d, 6, per=3, hist=False)
    assert rc == 0 and done == 6
    for i in range(3):
        np.testing.assert_array_equal(d.get_state(i), ref.get_state(i))


def test_history_beyond_the_scratch_is_refused():
    import turns2d_urans_git_amd as t2d
    from turns2d_urans_git_amd import abi
    d = H.gpu_case(H.cgrid(65, 25, ds_wall=1e-5), SA)
    n = (1 << 20) + 1
    h = np.zeros(1)   # never written: the call is refused before any work
    done = C.c_int32(-1)
    rc = _fn()(d.h, n, h.ctypes.data, C.byref(done))
    assert rc == abi.ERR_ARG
    assert "OT2D_NOALLOC_HIST_MAX" in t2d.last_error()
    assert d.istep0 == 0
