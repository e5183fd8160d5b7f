"""The request loop on the GPU: a result is the same alone and in a mix of
concurrently resident cases (bit for bit), admission keeps the resident
cases within the memory budget, and a converged case stops on the device."""
import numpy as np
import pytest

from turns2d_urans_git_amd import abi, grid, serve

pytestmark = pytest.mark.gpu


def _lib():
    import turns2d_urans_git_amd as t2d
    return t2d.load_library(), t2d.last_error


def _polar(rid, arrival, m=0.02, alfa=2.0, jmax=41, kmax=21, fp32=1):
    g = grid.naca4_cgrid(jmax, kmax, m=m, p=0.4, ds_wall=2e-4)
    deck = dict(fsmach=0.5, alfa=alfa, rey=1e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3, ilim=1,
                ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0,
                lu_fp32=fp32)
    return serve.Request(rid, arrival, deck, g, "polar")


def _pitch(rid, arrival, n=3):
    g = grid.naca4_cgrid(41, 21, ds_wall=2e-4)
    deck = dict(fsmach=0.3, alfa=-1.0, angmax=2.0, rf=0.05, rey=4e6, invisc=0, lamin=0, iturb=1, vnuinf=3.0,
                irhsy=-3, ilim=1, ilhs=1, iconstant_cfl=1, dt=5.0, epse=0.01, iunst=0, ntac=1, itnmax=1,
                timeac=1.0, timespectral=1, nspec=n, motiontype=1, lu_fp32=1)
    return serve.Request(rid, arrival, deck, g, "pitch")


TOL, K, MAXS = 1e-7, 50, 3000


def _serve(reqs, **kw):
    lib, err = _lib()
    srv = serve.Server(lib, steps_per_tick=K, tol=TOL, max_steps=MAXS, keep_state=True, errfn=err, **kw)
    return srv, srv.run(reqs)


@pytest.mark.parametrize("which", ["polar", "pitch"])
def test_same_alone_and_in_the_mix(which):
    target = _polar(0, 0.0, m=0.03, alfa=3.0) if which == "polar" else _pitch(0, 0.0, n=3)
    _, alone = _serve([target])
    mix = [_polar(1, 0.0, alfa=1.0), _polar(2, 0.0, jmax=49, alfa=-2.0), _pitch(3, 0.0, n=5),
           _polar(4, 0.0, fp32=0)]
    t2 = serve.Request(9, 0.05, target.deck, target.grid, target.kind)    # arrives while the mix runs
    _, both = _serve(mix + [t2, _polar(5, 0.1, alfa=5.0)])
    a, b = alone[0], [r for r in both if r.rid == 9][0]
    print(which, "alone:", a.status, a.steps, "mix:", b.status, b.steps, [(r.rid, r.status, r.steps) for r in both])
    assert a.status == b.status and a.steps == b.steps
    assert a.status in (abi.CONVERGED, -100)
    np.testing.assert_array_equal(a.state, b.state)
    np.testing.assert_array_equal(a.forces, b.forces)


def test_admission_respects_the_budget():
    lib, err = _lib()
    reqs = [_polar(i, 0.0, alfa=float(i % 4)) for i in range(6)]
    srv0 = serve.Server(lib, errfn=err)
    one = srv0.footprint(reqs[:1])
    srv, res = _serve(reqs, budget=int(2.5 * one))
    assert srv.peak_bytes <= int(2.5 * one)
    assert all(r.status in (abi.CONVERGED, -100) for r in res)
    adm = sorted(r.admitted for r in res)
    fin = sorted(r.finished for r in res)
    # at most two resident at a time: the third admission waits for a finish
    assert adm[2] >= fin[0]


def test_converged_case_stops_on_the_device():
    """With a tolerance the residual history reaches at step s, the case
    stops at exactly step s (status CONVERGED, before that step's update),
    and further iterations leave it untouched."""
    lib, err = _lib()
    r = _polar(0, 0.0)
    p = abi.Params()
    lib.ot2d_default_params.argtypes = [abi.C.POINTER(abi.Params)]
    lib.ot2d_default_params(abi.C.byref(p))
    p.update(**r.deck)
    md = abi.mesh_desc(r.grid.jmax, r.grid.kmax, grid.cgrid_bc(r.grid.jtail1))
    ref = abi.CaseAPI(lib, "ot2d_", [p], [md], errfn=err, batch=True)
    ref.set_grid(0, r.grid.xv, r.grid.yv)
    rc, hist, done = ref.iterate(600)
    assert done == 600
    tol = float(hist[300:, 0].min()) * (1 + 1e-12)
    s_first = int(np.argmax(hist[:, 0] <= tol)) + 1          # 1-based step
    case = abi.CaseAPI(lib, "ot2d_", [p], [md], errfn=err, batch=True)
    case.set_grid(0, r.grid.xv, r.grid.yv)
    case.set_convergence(tol)
    for _ in range(20):
        case.iterate_async(K)
        st, ist = case.batch_status()
        if st[0] == abi.CONVERGED:
            break
    assert st[0] == abi.CONVERGED and int(ist[0]) == s_first
    q1 = case.get_state(0)
    case.iterate_async(200)
    st2, ist2 = case.batch_status()
    assert st2[0] == abi.CONVERGED and int(ist2[0]) == s_first
    np.testing.assert_array_equal(case.get_state(0), q1)


def test_concurrent_cohorts_do_not_change_results():
    """Cases advanced on separate streams at the same time give the results
    they give alone, bit for bit (FP32 and FP64 LU-SGS, steady and
    time-spectral cohorts side by side).  These grids keep every FP32 sweep
    inside one workgroup; FP32 sweeps over several workgroups (the
    inter-workgroup mail) are covered by test_gpu_lu32_multiblock.py and by
    test_multiblock_fp32_cohorts_overlap below."""
    import turns2d_urans_git_amd as t2d
    lib, err = _lib()

    def batch(reqs):
        ps, mds = [], []
        for r in reqs:
            ps.append(t2d.default_params(**r.deck))
            mds.append(abi.mesh_desc(r.grid.jmax, r.grid.kmax, grid.cgrid_bc(r.grid.jtail1)))
        b = t2d.Batch(ps, mds)
        for i, r in enumerate(reqs):
            b.set_grid(i, r.grid.xv, r.grid.yv)
        return b

    target = _polar(0, 0.0, m=0.03, alfa=3.0)
    alone = batch([target])
    alone.iterate(600)
    ref = alone.get_state(0)
    alone.close()
    for trial in range(3):
        x = batch([target])
        others = [batch([_polar(1, 0.0, alfa=5.0), _polar(2, 0.0, jmax=49, alfa=2.0)]),
                  batch([_polar(3, 0.0, fp32=0)]), batch([_pitch(4, 0.0, n=3)])]
        for _ in range(12):
            x.iterate_async(50)
            for o in others:
                o.iterate_async(50)
            x.batch_status()
            for o in others:
                o.batch_status()
        np.testing.assert_array_equal(x.get_state(0), ref, err_msg=f"trial {trial}")
        x.close()
        for o in others:
            o.close()


def test_multiblock_fp32_cohorts_overlap():
    """Served cohorts whose FP32 sweeps span several workgroups (kmax 97 to
    161) beside other cohorts on their own streams: every request's result
    equals its result served alone, bit for bit."""
    target = _polar(0, 0.0, m=0.03, alfa=3.0, jmax=97, kmax=97)
    _, alone = _serve([target])
    mix = [_polar(1, 0.0, alfa=1.0, jmax=65, kmax=161), _polar(2, 0.0, jmax=81, kmax=129, alfa=-2.0),
           _pitch(3, 0.0, n=3), _polar(4, 0.0, fp32=0, kmax=97)]
    t2 = serve.Request(9, 0.05, target.deck, target.grid, target.kind)
    _, both = _serve(mix + [t2])
    a, b = alone[0], [r for r in both if r.rid == 9][0]
    assert a.status == b.status and a.steps == b.steps
    np.testing.assert_array_equal(a.state, b.state)
    np.testing.assert_array_equal(a.forces, b.forces)



def _batch_of(reqs, tol=None):
    import turns2d_urans_git_amd as t2d
    ps = [t2d.default_params(**r.deck) for r in reqs]
    mds = [abi.mesh_desc(r.grid.jmax, r.grid.kmax, grid.cgrid_bc(r.grid.jtail1)) for r in reqs]
    b = t2d.Batch(ps, mds)
    for i, r in enumerate(reqs):
        b.set_grid(i, r.grid.xv, r.grid.yv)
    if tol:
        b.set_convergence(tol)
    return b


@pytest.mark.parametrize("fp32", [1, 0])
def test_slot_admission_equals_alone(fp32):
    """A case put into the slot of a stopped one (ot2d_batch_admit: a smaller
    grid, other flow conditions) gives the state it gives alone after the
    same number of steps, bit for bit, and the case running beside it is
    not disturbed."""
    import turns2d_urans_git_amd as t2d
    first = _polar(0, 0.0, m=0.03, alfa=3.0, jmax=49, kmax=25, fp32=fp32)
    other = _polar(1, 0.0, alfa=1.0, jmax=41, kmax=21, fp32=fp32)
    new = _polar(2, 0.0, alfa=-2.0, m=0.01, jmax=45, kmax=21, fp32=fp32)
    n0, n1 = 200, 300
    ref = {}
    for r, n in ((other, n0 + n1), (new, n1)):
        b = _batch_of([r])
        rc, _, done = b.iterate(n)
        assert rc == 0 and done == n
        ref[r.rid] = (b.get_state(0), b.forces(0))
        b.close()
    b = _batch_of([first, other])
    rc, _, _ = b.iterate(n0)
    assert rc == 0
    b.batch_retire(0)
    b.batch_admit(0, t2d.default_params(**new.deck),
                  abi.mesh_desc(new.grid.jmax, new.grid.kmax, grid.cgrid_bc(new.grid.jtail1)))
    b.set_grid(0, new.grid.xv, new.grid.yv)
    rc, _, _ = b.iterate(n1)
    assert rc == 0
    st, ist = b.batch_status()
    assert list(st) == [0, 0] and list(ist) == [n1, n0 + n1]
    for i, r in ((0, new), (1, other)):
        np.testing.assert_array_equal(b.get_state(i), ref[r.rid][0], err_msg=f"case {r.rid}")
        assert b.forces(i) == ref[r.rid][1]
    b.close()


def test_slot_admission_refuses_what_does_not_fit():
    """A slot still running, a larger grid than the slot was made for, or
    another instance count: refused with an error."""
    import turns2d_urans_git_amd as t2d
    a = _polar(0, 0.0, jmax=41, kmax=21)
    b = _batch_of([a, _polar(1, 0.0, alfa=4.0, jmax=41, kmax=21)])
    big = _polar(2, 0.0, jmax=49, kmax=21)
    md_big = abi.mesh_desc(big.grid.jmax, big.grid.kmax, grid.cgrid_bc(big.grid.jtail1))
    md_a = abi.mesh_desc(a.grid.jmax, a.grid.kmax, grid.cgrid_bc(a.grid.jtail1))
    with pytest.raises(abi.CaseError):      # slot 0 still running
        b.batch_admit(0, t2d.default_params(**a.deck), md_a)
    b.batch_retire(0)
    with pytest.raises(abi.CaseError):      # 49 x 21 does not fit a 41 x 21 slot
        b.batch_admit(0, t2d.default_params(**big.deck), md_big)
    pitch = _pitch(3, 0.0, n=3)
    with pytest.raises(abi.CaseError):      # time-spectral in a steady batch
        b.batch_admit(0, t2d.default_params(**pitch.deck), md_a)
    b.close()


def test_server_fills_freed_slots():
    """With memory for one cohort only, a request that arrives while the
    first runs waits, then takes the first's slot when that stops (no new
    cohort), with the result it gives alone."""
    first = _polar(0, 0.0, m=0.03, alfa=3.0, jmax=49, kmax=25)
    second = _polar(1, 0.001, alfa=-1.0, jmax=41, kmax=21)
    lib, err = _lib()

    def run(reqs, **kw):
        srv = serve.Server(lib, steps_per_tick=K, tol=TOL, max_steps=400, keep_state=True, errfn=err, **kw)
        return srv, srv.run(reqs)

    one = serve.Server(lib, errfn=err).footprint([first])
    srv, res = run([first, second], budget=int(1.2 * one))
    assert srv.cohorts_made == 1 and srv.slot_admissions == 1
    _, alone = run([second])
    a, b = alone[0], res[1]
    assert a.status == b.status and a.steps == b.steps
    np.testing.assert_array_equal(a.state, b.state)
    np.testing.assert_array_equal(a.forces, b.forces)


def test_presized_cohorts_take_later_arrivals():
    """With slot_dims every cohort has max_cohort slots of its structure's
    largest grid: requests arriving while it runs take its free slots (FP32
    sweeps of W = 1 and W = 2 blocks side by side), a second cohort of the
    structure is made only when the first is full, and each result equals
    the request run alone, bit for bit."""
    lib, err = _lib()
    dims = {"polar": (57, 45), "pitch": (41, 21)}
    reqs = [_polar(0, 0.0, alfa=1.0), _polar(1, 0.0, jmax=49, kmax=41, alfa=-2.0), _pitch(2, 0.0, n=3),
            _polar(3, 0.02, m=0.03, alfa=3.0, kmax=37), _polar(4, 0.04, alfa=5.0, jmax=57),
            _polar(5, 0.06, alfa=0.5, kmax=45), _pitch(6, 0.08, n=3)]
    srv = serve.Server(lib, steps_per_tick=K, tol=TOL, max_steps=400, keep_state=True, errfn=err, max_cohort=3,
                       slot_dims=lambda r: dims[r.kind])
    res = srv.run(reqs)
    print("cohorts", srv.cohorts_made, "slot admissions", srv.slot_admissions,
          [(r.rid, r.status, r.steps) for r in res])
    # polar 0, 1 and pitch 2 start two cohorts; 3 fills the polar cohort's
    # third slot, 6 the pitch cohort's second; 4 and 5 find the polar cohort
    # full unless a case has finished
    assert srv.slot_admissions >= 2
    assert srv.cohorts_made <= 3
    for r in res:
        srv1 = serve.Server(lib, steps_per_tick=K, tol=TOL, max_steps=400, keep_state=True, errfn=err)
        a = srv1.run([serve.Request(r.rid, 0.0, reqs[r.rid].deck, reqs[r.rid].grid, reqs[r.rid].kind)])[0]
        assert a.status == r.status and a.steps == r.steps, r.rid
        assert a.status in (abi.CONVERGED, -100)
        np.testing.assert_array_equal(a.state, r.state, err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(a.forces, r.forces, err_msg=f"request {r.rid}")
