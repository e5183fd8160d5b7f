"""Host side of the request loop: the synthetic request streams and the
footprint used for admission (pure host code in libot2d, no GPU)."""
import os

import numpy as np
import pytest

from turns2d_urans_git_amd import abi, grid, serve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "turns2d_urans_git_amd", "libot2d.so")


def test_poisson_stream_mix():
    reqs = serve.poisson_stream(20.0, 30.0, seed=7, scale=0.25)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and arr[-1] <= 30.0
    assert 450 < len(reqs) < 750                      # 600 expected
    frac = np.mean([r.kind == "pitch" for r in reqs])
    assert 0.22 < frac < 0.38
    for r in reqs:
        if r.kind == "pitch":
            assert r.nspec in (3, 5, 7) and r.deck["timespectral"] == 1
            assert 1.0 <= r.deck["angmax"] <= 5.0 and 0.012 <= r.deck["rf"] <= 0.06 + 1e-12
        else:
            assert r.nspec == 1 and -4 <= r.deck["alfa"] <= 8
    # the same seed gives the same stream
    again = serve.poisson_stream(20.0, 30.0, seed=7, scale=0.25)
    assert [r.arrival for r in again] == list(arr)


def test_summary_percentiles():
    res = [serve.Result(i, "polar", 0.0, finished=float(i + 1), status=abi.CONVERGED, steps=10, cells=100)
           for i in range(100)]
    s = serve.summarize(res, 100.0)
    assert s["converged"] == 100 and abs(s["p50_s"] - 50.5) < 1e-9 and abs(s["p99_s"] - 99.01) < 1e-9
    assert abs(s["cell_iter_per_s"] - 100 * 10 * 100 / 100.0) < 1e-9


@pytest.mark.skipif(not os.path.exists(LIB), reason="libot2d.so not built")
def test_footprint_grows_with_the_mesh_and_instances():
    import turns2d_urans_git_amd as t2d
    lib = t2d.load_library()
    def fp(jmax, kmax, **kw):
        p = t2d.default_params(invisc=0, lamin=0, iturb=1, **kw)
        g = grid.naca4_cgrid(jmax, kmax)
        return abi.footprint(lib, "ot2d_", [p], [abi.mesh_desc(jmax, kmax, grid.cgrid_bc(g.jtail1))])
    a, b = fp(129, 49), fp(257, 97)
    assert b > 3.5 * a
    cells = 257 * 97
    assert 200 * cells < b < 2000 * cells            # bytes per cell of an SA case
    ts3 = fp(129, 49, timespectral=1, nspec=3, motiontype=1, rf=0.1)
    assert 2.5 * a < ts3 < 3.5 * a
