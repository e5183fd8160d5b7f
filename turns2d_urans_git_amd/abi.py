"""ctypes mirror of ``include/ot2d.h`` and a thin host-side case wrapper.

``CaseAPI`` drives any library that exports the ``ot2d.h`` entry points under
a given prefix.  The product library ``libot2d.so`` uses the prefix
``ot2d_``; the wrapper is deliberately free of torch types so the same host
code can feed NumPy arrays straight through the C-ABI.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Sequence

import numpy as np

MAXBC = 25

OK, MSTOP, NORM, CONVERGED = 0, 1, 2, 3
ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_STATE, ERR_NOMEM = -1, -2, -3, -4, -5

(F_X, F_Y, F_XX, F_XY, F_YX, F_YY, F_UG, F_VG, F_TSCALE, F_BT, F_TURMU, F_S,
 F_XV, F_YV, F_UGV, F_VGV) = range(16)


class BCSeg(C.Structure):
    _fields_ = [("ibtyp", C.c_int32), ("ibdir", C.c_int32), ("jbcs", C.c_int32),
                ("jbce", C.c_int32), ("kbcs", C.c_int32), ("kbce", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("jmax", C.c_int32), ("kmax", C.c_int32), ("nbc", C.c_int32),
                ("irot_dir", C.c_int32), ("xcen", C.c_double), ("ycen", C.c_double),
                ("bc", BCSeg * MAXBC)]


_PARAM_INTS = ["iread", "nhalo", "half", "invisc", "lamin", "iturb", "itrans",
               "irhsy", "ilim", "ilhs", "iconstant_cfl", "idual", "iprecon",
               "ntac", "iunst", "itnmax", "timespectral", "nspec", "motiontype",
               "nsteps", "npnorm", "nrest"]
_PARAM_DBLS = ["fsmach", "alfa", "rey", "vnuinf", "tuinf", "dt", "timeac", "epse",
               "dualtime", "Mp", "rf", "angmax", "fmtip", "rartio", "totime",
               "gamma", "pr", "rmue", "tinf"]


class Params(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _PARAM_INTS]
                + [(n, C.c_double) for n in _PARAM_DBLS]
                + [("lu_fp32", C.c_int32), ("static_conn", C.c_int32), ("reserved", C.c_int32 * 6)])

    def update(self, **kw):
        for k, v in kw.items():
            if k in _PARAM_INTS or k in ("lu_fp32", "static_conn"):
                setattr(self, k, int(bool(v)) if isinstance(v, bool) else int(v))
            elif k in _PARAM_DBLS:
                setattr(self, k, float(v))
            else:
                raise KeyError(f"unknown parameter {k}")
        return self

    def as_dict(self):
        return {n: getattr(self, n) for n in _PARAM_INTS + _PARAM_DBLS + ["lu_fp32"]}


class IHCBC(C.Structure):
    """One ``ihcbc.<n>`` file (``ot2d_ihcbc``)."""
    _fields_ = [("nbc", C.c_int32)] + [(n, C.c_int32 * MAXBC) for n in
                                        ("ibtyp", "ibdir", "jbcs", "jbce", "kbcs", "kbce")]


def ihcbc(segs: Sequence) -> IHCBC:
    b = IHCBC()
    b.nbc = len(segs)
    for i, s in enumerate(segs):
        for n in ("ibtyp", "ibdir", "jbcs", "jbce", "kbcs", "kbce"):
            getattr(b, n)[i] = getattr(s, n)
    return b


class IHCMesh(C.Structure):
    _fields_ = [("jmax", C.c_int32), ("kmax", C.c_int32),
                ("xv", C.c_void_p), ("yv", C.c_void_p), ("xc", C.c_void_p), ("yc", C.c_void_p),
                ("vol", C.c_void_p), ("bc", IHCBC)]


def mesh_desc(jmax: int, kmax: int, segs: Sequence, irot_dir: int = 1,
              xcen: float = 0.0, ycen: float = 0.0) -> MeshDesc:
    md = MeshDesc()
    md.jmax, md.kmax, md.nbc = jmax, kmax, len(segs)
    md.irot_dir, md.xcen, md.ycen = irot_dir, xcen, ycen
    for i, s in enumerate(segs):
        md.bc[i] = BCSeg(s.ibtyp, s.ibdir, s.jbcs, s.jbce, s.kbcs, s.kbce)
    return md


_DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_IP = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def bind(lib: C.CDLL, prefix: str):
    """Declare the ot2d.h signatures on ``lib`` under ``prefix``."""
    P = C.POINTER
    sig = {
        "default_params": (C.c_int32, [P(Params)]),
        "sizes": (C.c_int32, [P(Params), P(MeshDesc), P(C.c_int32), P(C.c_int32),
                              P(C.c_int32), P(C.c_int32)]),
        "case_create": (C.c_int32, [P(Params), C.c_int32, P(MeshDesc), P(C.c_void_p)]),
        "case_destroy": (C.c_int32, [C.c_void_p]),
        "set_grid": (C.c_int32, [C.c_void_p, C.c_int32, _DP, _DP]),
        "set_state": (C.c_int32, [C.c_void_p, C.c_int32, _DP]),
        "get_state": (C.c_int32, [C.c_void_p, C.c_int32, _DP]),
        "set_iblank": (C.c_int32, [C.c_void_p, C.c_int32, _IP]),
        "get_iblank": (C.c_int32, [C.c_void_p, C.c_int32, _IP]),
        "get_field": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _DP]),
        "set_field": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _DP]),
        "set_istep0": (C.c_int32, [C.c_void_p, C.c_int32]),
        "get_istep0": (C.c_int32, [C.c_void_p, P(C.c_int32)]),
        "time_step": (C.c_int32, [C.c_void_p]),
        "bc": (C.c_int32, [C.c_void_p]),
        "compute_residual": (C.c_int32, [C.c_void_p]),
        "rbc": (C.c_int32, [C.c_void_p]),
        "lhs": (C.c_int32, [C.c_void_p]),
        "turbulence": (C.c_int32, [C.c_void_p]),
        "monitor": (C.c_int32, [C.c_void_p, P(C.c_int32)]),
        "step": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "iterate": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, P(C.c_int32)]),
        "forces": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _DP]),
    }
    # present in libot2d only (no counterpart in the reference)
    optional = {
        "profile": (C.c_int32, [C.c_void_p, C.c_int32, _DP, P(C.c_int32)]),
        "stage_name": (C.c_char_p, [C.c_int32]),
        "batch_create": (C.c_int32, [C.c_int32, C.c_void_p, C.c_void_p, P(C.c_void_p)]),
        "batch_status": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "batch_retire": (C.c_int32, [C.c_void_p, C.c_int32]),
        "batch_admit": (C.c_int32, [C.c_void_p, C.c_int32, P(Params), P(MeshDesc)]),
        "batch_compact": (C.c_int32, [C.c_void_p, P(C.c_int32)]),
        "set_dtpseudo": (C.c_int32, [C.c_void_p, C.c_int32, _DP]),
        "subiter_norms": (C.c_int32, [C.c_void_p, _DP, _IP]),
        "set_launch": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
        "get_launch": (C.c_int32, [C.c_void_p, P(C.c_int32), P(C.c_int32)]),
        "autotune": (C.c_int32, [C.c_void_p, C.c_int32, P(C.c_int32)]),
        "connect": (C.c_int32, [C.c_void_p, C.c_void_p, P(C.c_int32)]),
        "interpolate": (C.c_int32, [C.c_void_p]),
        "set_connectivity": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
        "ihc_connect": (C.c_int32, [C.c_void_p, C.c_int32, P(C.c_void_p)]),
        "ihc_counts": (C.c_int32, [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                   P(C.c_int32)]),
        "ihc_get": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
        "ihc_free": (C.c_int32, [C.c_void_p]),
        "footprint": (C.c_int32, [C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, P(C.c_int64)]),
        "case_bytes": (C.c_int32, [C.c_void_p, P(C.c_int64)]),
        "set_convergence": (C.c_int32, [C.c_void_p, C.c_int32, C.c_double]),
        "iterate_async": (C.c_int32, [C.c_void_p, C.c_int32]),
        "sync": (C.c_int32, [C.c_void_p]),
        "set_fp32_order": (C.c_int32, [C.c_int32]),
        "query": (C.c_int32, [C.c_void_p, P(C.c_int32)]),
        # oracle only: connectivity of one mesh as plain arrays
        "set_mesh_connectivity": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _IP, C.c_int32, _IP, _IP,
                                              C.c_int32, _IP, _DP]),
    }
    fns = {}
    for name, (res, args) in list(sig.items()) + list(optional.items()):
        if name in optional and not hasattr(lib, prefix + name):
            continue
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    return fns


NSTAGE = 21  # OT2D_NSTAGE


class CaseError(RuntimeError):
    def __init__(self, rc, where, msg=""):
        super().__init__(f"{where} failed with status {rc}{': ' + msg if msg else ''}")
        self.rc = rc


class CaseAPI:
    """One case (``nmesh`` meshes x ``nspec`` instances) behind a C-ABI library."""

    def __init__(self, lib: C.CDLL, prefix: str, params, meshes: Sequence[MeshDesc],
                 errfn=None, batch: bool = False):
        """``params`` is one Params (a case of ``len(meshes)`` meshes) or, with
        ``batch=True``, one Params per mesh (independent cases, ot2d_batch_create)."""
        self._fns = bind(lib, prefix)
        self._errfn = errfn
        self.meshes = list(meshes)
        arr = (MeshDesc * len(meshes))(*meshes)
        self._arr = arr
        h = C.c_void_p()
        if batch:
            plist = list(params)
            assert len(plist) == len(meshes)
            parr = (Params * len(plist))(*plist)
            self._parr = parr
            self._check(self._fns["batch_create"](len(plist), C.cast(parr, C.c_void_p),
                                                  C.cast(arr, C.c_void_p), C.byref(h)), "batch_create")
            self.params = plist[0]
            per_mesh = plist
        else:
            self._check(self._fns["case_create"](C.byref(params), len(meshes), arr, C.byref(h)),
                        "case_create")
            self.params = params
            per_mesh = [params] * len(meshes)
        self.h = h
        self.batch = batch
        self.dims = []
        for md, pm in zip(meshes, per_mesh):
            jd, kd, nq, nv = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            self._fns["sizes"](C.byref(pm), C.byref(md), C.byref(jd), C.byref(kd),
                               C.byref(nq), C.byref(nv))
            self.dims.append((jd.value, kd.value, nq.value, nv.value, md.jmax, md.kmax))
        self.nspec = self.params.nspec if self.params.timespectral else 1

    # -------------------------------------------------------------- utils
    def _check(self, rc, where):
        if rc < 0:
            msg = self._errfn() if self._errfn else ""
            raise CaseError(rc, where, msg)
        return rc

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._fns["case_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- overset
    def connect(self, ihc_rows: Sequence[Sequence]) -> int:
        """Hole cutting and donor search over the case's meshes (one list of
        ihcbc rows per mesh); returns the fringe cells of all meshes."""
        arr = (IHCBC * len(ihc_rows))(*[ihcbc(r) for r in ihc_rows])
        n = C.c_int32()
        self._check(self._fns["connect"](self.h, arr, C.byref(n)), "connect")
        return n.value

    def set_connectivity(self, nsp: int, lists: "IHCLists"):
        self._check(self._fns["set_connectivity"](self.h, nsp, lists.h), "set_connectivity")

    def set_mesh_connectivity(self, nsp: int, im: int, iblank, imesh, ibc, idonor, frac):
        """Oracle form: one mesh's iblank and lists in the reference layout."""
        iblank = np.ascontiguousarray(np.asarray(iblank).ravel(order="F"), dtype=np.int32)
        imesh = np.ascontiguousarray(imesh, dtype=np.int32)
        ibc = np.ascontiguousarray(ibc, dtype=np.int32)
        idonor = np.ascontiguousarray(idonor, dtype=np.int32)
        frac = np.ascontiguousarray(frac, dtype=np.float64)
        self._check(self._fns["set_mesh_connectivity"](self.h, nsp, im, iblank, len(ibc), imesh, ibc,
                                                       len(idonor) // 2, idonor, frac),
                    "set_mesh_connectivity")

    # ---------------------------------------------------------- serving
    def set_convergence(self, tol: float, icase: int = -1):
        self._check(self._fns["set_convergence"](self.h, icase, float(tol)), "set_convergence")

    def iterate_async(self, niter: int):
        self._check(self._fns["iterate_async"](self.h, int(niter)), "iterate_async")

    def idle(self) -> bool:
        """True when all work enqueued for the case has finished (ot2d_query;
        never waits)."""
        v = C.c_int32()
        self._check(self._fns["query"](self.h, C.byref(v)), "query")
        return bool(v.value)

    def sync(self):
        self._check(self._fns["sync"](self.h), "sync")

    def bytes(self) -> int:
        b = C.c_int64()
        self._check(self._fns["case_bytes"](self.h, C.byref(b)), "case_bytes")
        return b.value

    def interpolate(self):
        self._check(self._fns["interpolate"](self.h), "interpolate")

    # ------------------------------------------------------------ state
    def set_grid(self, im: int, xv: np.ndarray, yv: np.ndarray):
        xv = np.ascontiguousarray(np.asfortranarray(xv).ravel(order="F"))
        yv = np.ascontiguousarray(np.asfortranarray(yv).ravel(order="F"))
        self._check(self._fns["set_grid"](self.h, im, xv, yv), "set_grid")

    def qshape(self, im):
        jd, kd, nq, nv, _, _ = self.dims[im]
        return (jd, kd, self.nspec, nq)

    def get_state(self, im: int = 0) -> np.ndarray:
        jd, kd, nq, nv, _, _ = self.dims[im]
        buf = np.empty(jd * kd * self.nspec * nq)
        self._check(self._fns["get_state"](self.h, im, buf), "get_state")
        return buf.reshape((jd, kd, self.nspec, nq), order="F")

    def set_state(self, im: int, q: np.ndarray):
        buf = np.ascontiguousarray(np.asarray(q, dtype=np.float64).ravel(order="F"))
        self._check(self._fns["set_state"](self.h, im, buf), "set_state")

    def get_iblank(self, im: int = 0) -> np.ndarray:
        jd, kd, *_ = self.dims[im]
        buf = np.empty(jd * kd * self.nspec, dtype=np.int32)
        self._check(self._fns["get_iblank"](self.h, im, buf), "get_iblank")
        return buf.reshape((jd, kd, self.nspec), order="F")

    def set_iblank(self, im: int, ib: np.ndarray):
        buf = np.ascontiguousarray(np.asarray(ib, dtype=np.int32).ravel(order="F"))
        self._check(self._fns["set_iblank"](self.h, im, buf), "set_iblank")

    def _field_shape(self, im, field):
        jd, kd, nq, nv, jmax, kmax = self.dims[im]
        if field == F_S:
            return (jd, kd, self.nspec, nv)
        if field in (F_XV, F_YV, F_UGV, F_VGV):
            return (jmax, kmax, self.nspec)
        return (jd, kd, self.nspec)

    def get_field(self, im: int, field: int) -> np.ndarray:
        shp = self._field_shape(im, field)
        buf = np.empty(int(np.prod(shp)))
        self._check(self._fns["get_field"](self.h, im, field, buf), "get_field")
        return buf.reshape(shp, order="F")

    def set_field(self, im: int, field: int, a: np.ndarray):
        buf = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        self._check(self._fns["set_field"](self.h, im, field, buf), "set_field")

    @property
    def istep0(self) -> int:
        v = C.c_int32()
        self._fns["get_istep0"](self.h, C.byref(v))
        return v.value

    @istep0.setter
    def istep0(self, v: int):
        self._fns["set_istep0"](self.h, int(v))

    # ------------------------------------------------------------ stages
    def time_step(self):
        return self._check(self._fns["time_step"](self.h), "time_step")

    def bc(self):
        return self._check(self._fns["bc"](self.h), "bc")

    def compute_residual(self):
        return self._check(self._fns["compute_residual"](self.h), "compute_residual")

    def rbc(self):
        return self._check(self._fns["rbc"](self.h), "rbc")

    def lhs(self):
        return self._check(self._fns["lhs"](self.h), "lhs")

    def turbulence(self):
        return self._check(self._fns["turbulence"](self.h), "turbulence")

    def monitor(self) -> int:
        m = C.c_int32()
        self._check(self._fns["monitor"](self.h, C.byref(m)), "monitor")
        return m.value

    def step(self):
        n = len(self.meshes) * self.nspec
        norms = np.zeros(3 * n)
        am = np.zeros(3 * n, dtype=np.int32)
        rc = self._check(self._fns["step"](self.h, norms.ctypes.data, am.ctypes.data), "step")
        return rc, norms.reshape(n, 3), am.reshape(n, 3)

    def iterate(self, niter: int):
        n = len(self.meshes) * self.nspec
        hist = np.zeros(niter * n)
        done = C.c_int32()
        rc = self._check(self._fns["iterate"](self.h, int(niter), hist.ctypes.data,
                                              C.byref(done)), "iterate")
        return rc, hist.reshape(niter, n)[: done.value], done.value

    def profile(self, nsteps: int):
        """Advance ``nsteps`` eager steps timing every kernel stage with HIP
        events on the case's stream; returns (rc, {stage: mean ms/step}, done)."""
        ms = np.zeros(NSTAGE)
        done = C.c_int32()
        rc = self._check(self._fns["profile"](self.h, int(nsteps), ms, C.byref(done)), "profile")
        names = [self._fns["stage_name"](i).decode() for i in range(NSTAGE)]
        return rc, {n: float(v) for n, v in zip(names, ms) if v > 0.0}, done.value

    def batch_status(self):
        """(status per case, istep0 per case) of a batch."""
        n = len(self.meshes)
        st = np.zeros(n, dtype=np.int32)
        ist = np.zeros(n, dtype=np.int32)
        self._check(self._fns["batch_status"](self.h, st.ctypes.data, ist.ctypes.data), "batch_status")
        return st, ist

    def batch_retire(self, icase: int):
        return self._check(self._fns["batch_retire"](self.h, int(icase)), "batch_retire")

    def batch_admit(self, icase: int, params: "Params", mesh: "MeshDesc"):
        """Put a new case (params, mesh) into the stopped slot icase; its grid
        is set with set_grid next (ot2d_batch_admit)."""
        rc = self._check(self._fns["batch_admit"](self.h, int(icase), C.byref(params), C.byref(mesh)),
                         "batch_admit")
        self.meshes[icase] = mesh
        jd, kd, nq, nv = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self._fns["sizes"](C.byref(params), C.byref(mesh), C.byref(jd), C.byref(kd), C.byref(nq), C.byref(nv))
        self.dims[icase] = (jd.value, kd.value, nq.value, nv.value, mesh.jmax, mesh.kmax)
        return rc

    def set_dtpseudo(self, values):
        """dtpseudo(itn) of sub-iterations 1..len(values) (ot2d_set_dtpseudo)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        return self._check(self._fns["set_dtpseudo"](self.h, len(v), v), "set_dtpseudo")

    def subiter_norms(self, itnmax: int):
        """Norms (itnmax, instances, 3) and argmax of every sub-iteration of
        the last time step (ot2d_subiter_norms)."""
        per = len(self.meshes) * self.nspec
        n = np.zeros(itnmax * per * 3)
        a = np.zeros(itnmax * per * 3, dtype=np.int32)
        self._check(self._fns["subiter_norms"](self.h, n, a), "subiter_norms")
        return n.reshape(itnmax, per, 3), a.reshape(itnmax, per, 3)

    def batch_compact(self) -> int:
        """Launch the following steps over the running cases only
        (ot2d_batch_compact); returns the number of running cases."""
        n = C.c_int32()
        self._check(self._fns["batch_compact"](self.h, C.byref(n)), "batch_compact")
        return n.value

    def set_launch(self, sweep: int = 0, sa_lane: int = -1):
        """Launch choices that change speed, not results (ot2d_set_launch)."""
        return self._check(self._fns["set_launch"](self.h, int(sweep), int(sa_lane)), "set_launch")

    def get_launch(self):
        """(sweep, sa_lane) in use (ot2d_get_launch)."""
        sw, sl = C.c_int32(), C.c_int32()
        self._check(self._fns["get_launch"](self.h, C.byref(sw), C.byref(sl)), "get_launch")
        return sw.value, sl.value

    def autotune(self, steps: int = 10):
        """Measured search over the launch choices for this case's shape
        class, cached per class (ot2d_autotune).  Returns (sweep, sa_lane,
        measured); measured is False on a cache hit."""
        m = C.c_int32()
        self._check(self._fns["autotune"](self.h, int(steps), C.byref(m)), "autotune")
        sw, sl = self.get_launch()
        return sw, sl, bool(m.value)

    def forces(self, im: int = 0, nsp: int = 0):
        out = np.zeros(6)
        self._check(self._fns["forces"](self.h, im, nsp, out), "forces")
        return dict(zip(("cl", "cd", "cm", "cpower", "cfx", "cfy"), out))


def default_params(lib: C.CDLL, prefix: str, **overrides) -> Params:
    p = Params()
    f = getattr(lib, prefix + "default_params")
    f.restype = C.c_int32
    f.argtypes = [C.POINTER(Params)]
    f(C.byref(p))
    p.update(**overrides)
    return p


# ------------------------------------------------------------------ overset
class IHCLists:
    """Result of ``ot2d_ihc_connect`` (owned; freed with the object)."""

    def __init__(self, lib, prefix, handle, dims, errfn=None):
        self._fns = bind(lib, prefix)
        self.h = handle
        self.dims = dims
        self._err = errfn

    def __del__(self):
        if getattr(self, "h", None):
            self._fns["ihc_free"](self.h)
            self.h = None

    def counts(self, im):
        a = [C.c_int32() for _ in range(4)]
        rc = self._fns["ihc_counts"](self.h, im, *[C.byref(x) for x in a])
        if rc:
            raise CaseError(rc, "ihc_counts", self._err() if self._err else "")
        return tuple(x.value for x in a)

    def get(self, im):
        """iblank (jmax+1, kmax+1) and the lists of mesh im, as the
        reference lays them out."""
        jmax, kmax = self.dims[im]
        nf, nd, _, _ = self.counts(im)
        ib = np.zeros((jmax + 1) * (kmax + 1), dtype=np.int32)
        imesh = np.zeros(2 * nf, dtype=np.int32)
        ibc = np.zeros(nf, dtype=np.int32)
        idonor = np.zeros(2 * nd, dtype=np.int32)
        frac = np.zeros(2 * nd, dtype=np.float64)
        rc = self._fns["ihc_get"](self.h, im, ib.ctypes.data, imesh.ctypes.data, ibc.ctypes.data,
                                  idonor.ctypes.data, frac.ctypes.data)
        if rc:
            raise CaseError(rc, "ihc_get", self._err() if self._err else "")
        return dict(iblank=ib, imesh=imesh, ibc=ibc, idonor=idonor, frac=frac)


def ihc_connect(lib, prefix, meshes, errfn=None) -> IHCLists:
    """Hole cutting and donor search over ``meshes``: dicts with ``xv, yv``
    (jmax, kmax), ``xc, yc, vol`` ((jmax+1, kmax+1), halo included) and
    ``ihc`` (rows of the mesh's ihcbc file)."""
    fns = bind(lib, prefix)
    arr = (IHCMesh * len(meshes))()
    keep = []
    dims = []
    for i, m in enumerate(meshes):
        xv, yv = (np.asfortranarray(m[k], dtype=np.float64) for k in ("xv", "yv"))
        xc, yc, vol = (np.asfortranarray(m[k], dtype=np.float64) for k in ("xc", "yc", "vol"))
        jmax, kmax = xv.shape
        assert xc.shape == (jmax + 1, kmax + 1), (xc.shape, jmax, kmax)
        keep += [xv, yv, xc, yc, vol]
        arr[i].jmax, arr[i].kmax = jmax, kmax
        arr[i].xv, arr[i].yv = xv.ctypes.data, yv.ctypes.data
        arr[i].xc, arr[i].yc, arr[i].vol = xc.ctypes.data, yc.ctypes.data, vol.ctypes.data
        arr[i].bc = ihcbc(m.get("ihc", []))
        dims.append((jmax, kmax))
    h = C.c_void_p()
    rc = fns["ihc_connect"](arr, len(meshes), C.byref(h))
    if rc:
        raise CaseError(rc, "ihc_connect", errfn() if errfn else "")
    return IHCLists(lib, prefix, h, dims, errfn)


def set_fp32_order(lib, prefix, on: bool):
    """Order FP32 LU-SGS work of different cases across streams (on, the
    default) or let it overlap (off), for the whole process
    (ot2d_set_fp32_order)."""
    rc = bind(lib, prefix)["set_fp32_order"](1 if on else 0)
    if rc:
        raise CaseError(rc, "set_fp32_order")


def footprint(lib, prefix, params_list, meshes) -> int:
    """Device bytes of a batch of these cases (or one case when a single
    Params is given for several meshes)."""
    fns = bind(lib, prefix)
    plist = list(params_list) if isinstance(params_list, (list, tuple)) else [params_list]
    parr = (Params * len(plist))(*plist)
    marr = (MeshDesc * len(meshes))(*meshes)
    b = C.c_int64()
    rc = fns["footprint"](len(plist), C.cast(parr, C.c_void_p), len(meshes), C.cast(marr, C.c_void_p), C.byref(b))
    if rc:
        raise CaseError(rc, "footprint")
    return b.value
