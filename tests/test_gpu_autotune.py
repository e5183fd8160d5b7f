"""Launch choices and the measured search over them (ot2d_set_launch,
ot2d_autotune).

The choices (the FP32 sweep's LDS ring depth and producer count, the SA line
solves' mapping) change only how the same operations are spread over lanes,
waves and workgroups, never their order within a cell or a line, so every
choice must give bit-identical states:

* each candidate, 20 steps of an SA FP32 case from the free stream: state,
  eddy viscosity and residual history equal those of the default choice;
* a search run mid-trajectory leaves the case where it was: after it, the
  trajectory equals that of an untuned twin, bit for bit, with the same
  step counter;
* a batch of mixed cases searched the same way equals its untuned twin;
* a second case of the same shape class takes the cached choice without
  measuring, and a class written to the OT2D_TUNE_CACHE file by one process
  is read by the next.
"""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from tests import helpers as H
from turns2d_urans_git_amd import abi, grid

pytestmark = pytest.mark.gpu

SA_DECK = dict(fsmach=0.7, alfa=2.0, rey=9.0e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3,
               ilim=1, ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1,
               timeac=1.0)

SWEEPS = (0, 1, 2)
SA_LANES = (-1, 0, 1)


def _snap(c):
    return c.get_state(0), c.get_field(0, abi.F_TURMU)


def test_every_launch_choice_is_bit_identical():
    g = H.cgrid(257, 97, ds_wall=1e-5)
    print("[autotune] choices 257x97 SA fp32", flush=True)
    ref = None
    for sw in SWEEPS:
        for sl in SA_LANES:
            c = H.gpu_case(g, dict(SA_DECK, lu_fp32=1))
            c.set_launch(sw, sl)
            assert c.get_launch() == (sw, sl)
            rc, hist, done = c.iterate(20)
            assert rc == 0 and done == 20
            got = (hist.copy(),) + _snap(c)
            c.close()
            if ref is None:
                ref = got
                continue
            for a, b, what in zip(got, ref, ("history", "q", "turmu")):
                np.testing.assert_array_equal(a, b, err_msg=f"{what}, sweep={sw} sa_lane={sl}")


def test_bad_choice_is_refused():
    c = H.gpu_case(H.cgrid(65, 25), dict(SA_DECK, lu_fp32=1))
    for sw, sl in ((-1, 0), (3, 0), (0, 2), (0, -2)):
        with pytest.raises(abi.CaseError):
            c.set_launch(sw, sl)
    assert c.get_launch() == (0, -1)
    c.close()


def test_search_leaves_the_trajectory_unchanged():
    g = H.cgrid(257, 81, ds_wall=1e-5)
    print("[autotune] search mid-trajectory 257x81", flush=True)
    tuned = H.gpu_case(g, dict(SA_DECK, lu_fp32=1))
    plain = H.gpu_case(g, dict(SA_DECK, lu_fp32=1))
    for c in (tuned, plain):
        rc, _, _ = c.iterate(7)
        assert rc == 0
    sw, sl, measured = tuned.autotune(steps=4)
    assert measured
    assert 0 <= sw < 3 and sl in (0, 1)
    assert tuned.istep0 == plain.istep0
    for a, b in zip(_snap(tuned), _snap(plain)):
        np.testing.assert_array_equal(a, b, err_msg="state right after the search")
    r1, h1, _ = tuned.iterate(15)
    r2, h2, _ = plain.iterate(15)
    assert r1 == r2 == 0
    np.testing.assert_array_equal(h1, h2)
    for a, b in zip(_snap(tuned), _snap(plain)):
        np.testing.assert_array_equal(a, b, err_msg="state 15 steps after the search")
    # a second case of the same class takes the cached choice
    other = H.gpu_case(g, dict(SA_DECK, lu_fp32=1, fsmach=0.5))
    assert other.autotune(steps=4) == (sw, sl, False)
    for c in (tuned, plain, other):
        c.close()


def test_batch_search_leaves_the_batch_unchanged():
    import turns2d_urans_git_amd as t2d
    shapes = [(129, 49), (161, 65), (193, 97), (129, 73), (257, 65)]
    grids = [H.cgrid(j, k, ds_wall=1e-5) for j, k in shapes]
    params = [t2d.default_params(**dict(SA_DECK, lu_fp32=1, alfa=1.0 + i)) for i in range(len(grids))]
    meshes = [abi.mesh_desc(g.jmax, g.kmax, grid.cgrid_bc(g.jtail1)) for g in grids]

    def make():
        b = t2d.Batch(params, meshes)
        for i, g in enumerate(grids):
            b.set_grid(i, g.xv, g.yv)
        return b

    tuned, plain = make(), make()
    print("[autotune] batch of 5", flush=True)
    for b in (tuned, plain):
        rc, _, _ = b.iterate(3)
        assert rc == 0
    tuned.autotune(steps=3)
    r1, h1, _ = tuned.iterate(12)
    r2, h2, _ = plain.iterate(12)
    assert r1 == r2 == 0
    np.testing.assert_array_equal(h1, h2)
    for i in range(len(grids)):
        np.testing.assert_array_equal(tuned.get_state(i), plain.get_state(i), err_msg=f"case {i}")
    tuned.close()
    plain.close()


_CHILD = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    from tests import helpers as H
    c = H.gpu_case(H.cgrid(97, 45, ds_wall=1e-5), {deck!r})
    print(c.autotune(steps=3), flush=True)
    c.close()
""")


def test_cache_file_is_shared_between_processes(tmp_path):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, OT2D_TUNE_CACHE=str(cache))
    code = _CHILD.format(root=root, deck=dict(SA_DECK, lu_fp32=1))
    outs = []
    for n in range(2):
        print(f"[autotune] cache file, process {n}", flush=True)
        r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True, text=True,
                           timeout=100)
        assert r.returncode == 0, r.stderr
        outs.append(eval(r.stdout.strip().splitlines()[-1]))
    assert outs[0][2] is True and outs[1][2] is False
    assert outs[0][:2] == outs[1][:2]
    lines = cache.read_text().split("\n")
    assert len([ln for ln in lines if ln.strip()]) == 1 and "fp32=1" in lines[0]
