"""Shared test helpers: build the same case on the oracle and on the GPU.

The oracle (oracle/liboracle.so) is test infrastructure; only tests,
__graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from turns2d_urans_git_amd import abi, grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_HARNESS = os.path.join(ORACLE_DIR, "_ref", "ref_harness")
REF_PROGRAM = os.path.join(ORACLE_DIR, "_ref", "overturns2d")

_oracle = None


def oracle_lib() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        _oracle = C.CDLL(ORACLE_LIB)
    return _oracle


# Euler deck of the C-grid airfoil cases (MUSCL kappa=1/3, Koren, Roe,
# LU-SGS at constant CFL): the namelist values the reference reads.
EULER = dict(fsmach=0.5, alfa=1.25, invisc=1, lamin=1, irhsy=-3, ilim=1, ilhs=1,
             iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0)


def cgrid(jmax=65, kmax=25, ds_wall=1e-3, m=0.0, p=0.0, t=0.12):
    return grid.naca4_cgrid(jmax, kmax, m=m, p=p, t=t, ds_wall=ds_wall)


def params_for(lib, prefix, **deck):
    p = abi.Params()
    f = getattr(lib, prefix + "default_params")
    f.argtypes = [C.POINTER(abi.Params)]
    f(C.byref(p))
    p.update(**deck)
    return p


def make_case(lib, prefix, g, deck, errfn=None, segs=None):
    p = params_for(lib, prefix, **deck)
    md = abi.mesh_desc(g.jmax, g.kmax, segs or grid.cgrid_bc(g.jtail1))
    case = abi.CaseAPI(lib, prefix, p, [md], errfn=errfn)
    case.set_grid(0, g.xv, g.yv)
    return case


def oracle_case(g, deck, segs=None):
    return make_case(oracle_lib(), "or_", g, deck, segs=segs)


def gpu_case(g, deck, segs=None):
    import turns2d_urans_git_amd as t2d
    lib = t2d.load_library()
    return make_case(lib, "ot2d_", g, deck, errfn=t2d.last_error, segs=segs)


def rel_err(a, b):
    """max |a-b| / max(|b|) (inf-norm relative error; 0 for identical arrays)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b).max() if a.size else 0.0
    s = np.abs(b).max() if b.size else 0.0
    return 0.0 if d == 0.0 else d / max(s, 1e-300)


def assert_cellwise(a, b, rtol, what=""):
    """Per-cell relative check with a floor proportional to the array's
    inf-norm: |a-b| <= rtol * (|b| + max|b|).  The floor keeps cells whose
    value cancels to ~0 (free-stream residuals) from dominating."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert np.all(np.isfinite(a)), f"{what}: non-finite values"
    scale = np.abs(b).max()
    bad = np.abs(a - b) > rtol * (np.abs(b) + scale)
    assert not bad.any(), (f"{what}: {bad.sum()} cells differ; max abs diff "
                          f"{np.abs(a - b).max():.3e}, scale {scale:.3e}")


# ---------------------------------------------------------------- fixtures
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    deck = {k[5:]: float(z[k]) for k in z.files if k.startswith("deck_")}
    ints = {"invisc", "lamin", "irhsy", "ilim", "ilhs", "iconstant_cfl", "iunst", "ntac",
            "itnmax", "iturb", "timespectral", "nspec", "motiontype"}
    deck = {k: (int(v) if k in ints else v) for k, v in deck.items()}
    g = grid.CGrid(xv=np.asfortranarray(z["xv"]), yv=np.asfortranarray(z["yv"]),
                   jtail1=int(z["jtail1"]), jmax=z["xv"].shape[0], kmax=z["xv"].shape[1])
    return z, g, deck


def golden_case(lib, prefix, name, errfn=None, **extra):
    z, g, deck = load_golden(name)
    deck = dict(deck, **extra)
    return z, make_case(lib, prefix, g, deck, errfn=errfn)


def replay_stages(case, z, ol=None):
    """Replay the stage sequence of the fixture's dump iteration on ``case``
    starting from the fixture's input state; returns {stage: array}."""
    jd, kd, nq, nv, _, _ = case.dims[0]
    sh_q, sh_s, sh_1 = (jd, kd, 1, nq), (jd, kd, 1, nv), (jd, kd, 1)
    case.set_state(0, z["qin"].reshape(sh_q, order="F"))
    if "turmuin" in z.files:
        case.set_field(0, abi.F_TURMU, z["turmuin"].reshape(sh_1, order="F"))
    case.istep0 = int(z["dumpat"])
    out = {}
    case.time_step()
    out["tscale"] = case.get_field(0, abi.F_TSCALE)
    case.set_field(0, abi.F_S, np.zeros(sh_s))
    case.compute_residual()
    out["qbc"] = case.get_state(0)
    out["res"] = case.get_field(0, abi.F_S)
    case.rbc()
    out["rbc"] = case.get_field(0, abi.F_S)
    case.lhs()
    out["dq"] = case.get_field(0, abi.F_S)
    out["qlu"] = case.get_state(0)
    if ol is not None:
        ol(case.h)
        out["qturb"] = case.get_state(0)
        out["turmu"] = case.get_field(0, abi.F_TURMU)
    return out
