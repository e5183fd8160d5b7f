"""Shared pieces of the overset tests: the fixtures written by the
reference's connect2d / do_interpolations (tests/golden/make_golden.py) and
cases built from them."""
from __future__ import annotations

import hashlib
import os

import numpy as np

from tests import helpers as H
from turns2d_urans_git_amd import abi, grid


def load(name):
    z = np.load(os.path.join(H.GOLDEN, name + ".npz"), allow_pickle=False)
    deck = {k[5:]: float(z[k]) for k in z.files if k.startswith("deck_")}
    ints = {"invisc", "lamin", "irhsy", "ilim", "ilhs", "iconstant_cfl", "iunst", "ntac", "itnmax", "iturb"}
    deck = {k: (int(v) if k in ints else v) for k, v in deck.items()}
    return z, deck


def rows(a):
    return [grid.IHCSeg(*[int(v) for v in r]) for r in np.asarray(a).reshape(-1, 6)]


def bcs(a):
    return [grid.BCSeg(*[int(v) for v in r]) for r in np.asarray(a).reshape(-1, 6)]


def meshes(z, near=None, back=None):
    """(xv, yv, bc rows, ihc rows) of both meshes of a fixture."""
    if "xv1" in z.files:
        return [(z[f"xv{i}"], z[f"yv{i}"], bcs(z[f"bc{i}"]), rows(z[f"ihc{i}"])) for i in (1, 2)]
    ms = grid.airfoil_overset(near=near, back=back, ds_wall=1e-4)
    out = []
    for i, m in enumerate(ms, 1):
        h = hashlib.sha256(np.ascontiguousarray(m.xv).tobytes() + np.ascontiguousarray(m.yv).tobytes())
        assert h.hexdigest() == str(z[f"grid_sha{i}"]), "grid generator no longer reproduces the fixture grid"
        out.append((m.xv, m.yv, bcs(z[f"bc{i}"]), rows(z[f"ihc{i}"])))
    return out


def make(lib, prefix, ms, deck, errfn=None):
    p = H.params_for(lib, prefix, **deck)
    mds = [abi.mesh_desc(xv.shape[0], xv.shape[1], bc) for xv, yv, bc, _ in ms]
    case = abi.CaseAPI(lib, prefix, p, mds, errfn=errfn)
    for im, (xv, yv, _, _) in enumerate(ms):
        case.set_grid(im, xv, yv)
    return case


def install_reference_connectivity(case, z):
    for im in (1, 2):
        case.set_mesh_connectivity(0, im - 1, z[f"iblank{im}"], z[f"imesh{im}"], z[f"ibc{im}"],
                                   z[f"idonor{im}"], z[f"frac{im}"])


def reference_norms(z, nsteps):
    ref = np.zeros((nsteps, 2))
    for row in z["norms"]:
        ref[int(row[0]) - 1, int(row[1]) - 1] = row[3]
    return ref
