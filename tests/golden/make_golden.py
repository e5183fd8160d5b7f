"""Generate the golden fixtures under tests/golden/ from the reference itself.

Runs oracle/_ref/ref_harness (the reference OVERTURNS-2D subroutines built
from their own sources by oracle/Makefile) on small C-grid decks written by
turns2d_urans_git_amd.grid and stores, per case, the grid, the deck, the
per-stage arrays of one iteration, the residual history, the forces and the
final state as a compressed .npz.  The fixtures are data only; this script is
how they were made:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from turns2d_urans_git_amd import grid  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
# komegasst reads F1, F2 and strain on the halo ring without writing them
# there (src/komegasst.f:190-215, 593-844); the reference is run with
# zero-filled allocations so those reads are defined (zero)
ZERO_MALLOC = {"GLIBC_TUNABLES": "glibc.malloc.perturb=255"}

EULER = dict(fsmach=0.5, alfa=1.25, invisc=True, lamin=True, irhsy=-3, ilim=1, ilhs=1,
             iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0)
SA = dict(fsmach=0.5, alfa=2.0, rey=1.0e6, invisc=False, lamin=False, iturb=1, vnuinf=3.0,
          irhsy=-3, ilim=1, ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1,
          itnmax=1, timeac=1.0)

CASES = {
    # name: (jmax, kmax, ds_wall, camber m, p, t, deck, nsteps, dump iteration)
    "euler_33x17": (33, 17, 1e-3, 0.0, 0.0, 0.12, EULER, 40, 20),
    "sa_41x21": (41, 21, 1e-4, 0.02, 0.4, 0.12, SA, 40, 35),
    # WENO5 reconstruction (ilim=4, src/eulerrhs.f:1011-1230)
    "weno_sa_41x21": (41, 21, 1e-4, 0.02, 0.4, 0.12, dict(SA, ilim=4), 40, 35),
    # k-omega SST (iturb=2, src/komegasst.f)
    "sst_41x21": (41, 21, 1e-4, 0.02, 0.4, 0.12, dict(SA, iturb=2), 40, 35),
}

STAGES = ["qin", "turmuin", "tscale", "qbc", "res", "rbc", "sdt", "dq", "qlu", "qturb", "qout", "turmu"]


def run_case(name, jmax, kmax, ds, m, p, t, deck, nsteps, dumpat):
    g = grid.naca4_cgrid(jmax, kmax, m=m, p=p, t=t, ds_wall=ds)
    with tempfile.TemporaryDirectory() as d:
        grid.write_case(d, g, nsteps=nsteps, nrest=10**6, npnorm=1, **deck)
        out = os.path.join(d, "out")
        os.makedirs(out)
        env = dict(os.environ, OT2D_OUT=out, OT2D_NSTEPS=str(nsteps), OT2D_DUMPAT=str(dumpat), **ZERO_MALLOC)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([HARNESS], cwd=d, stdin=fin, env=env, check=True,
                           stdout=subprocess.DEVNULL)
        arrays = {}
        for st in STAGES + ["qfinal", "turmufinal", "x", "y", "xx", "xy", "yx", "yy", "qinit"]:
            fn = os.path.join(out, f"{st}_m1.bin")
            if os.path.exists(fn):
                arrays[st] = np.fromfile(fn)
        norms = np.loadtxt(os.path.join(out, "norms.txt"), comments="#", ndmin=2)
        forces = np.loadtxt(os.path.join(out, "forces.txt"), comments="#", ndmin=2)
    deck_arr = {f"deck_{k}": np.asarray(float(v)) for k, v in deck.items()}
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        xv=g.xv, yv=g.yv, jtail1=np.asarray(g.jtail1), nsteps=np.asarray(nsteps),
        dumpat=np.asarray(dumpat), norms=norms, forces=forces, **arrays, **deck_arr)
    print(name, "norm[-1]", norms[-1, 3], "cl", forces[-1, 1])


# Time-spectral fixtures (pitching airfoil, N instances coupled by the
# spectral derivative): residual history of every instance and final state.
TS_CASES = {
    "ts3_euler_41x21": (41, 21, 1e-3, dict(EULER, fsmach=0.3, alfa=1.0, timespectral=True, nspec=3,
                                           motiontype=1, rf=0.1, angmax=2.0), 12),
    "ts4_sa_41x21": (41, 21, 2e-4, dict(SA, fsmach=0.3, alfa=1.0, rey=4.0e6, dt=5.0, timespectral=True,
                                        nspec=4, motiontype=1, rf=0.2, angmax=3.0), 12),
}


def run_ts_case(name, jmax, kmax, ds, deck, nsteps):
    g = grid.naca4_cgrid(jmax, kmax, ds_wall=ds)
    with tempfile.TemporaryDirectory() as d:
        grid.write_case(d, g, nsteps=nsteps, nrest=10**6, npnorm=1, **deck)
        out = os.path.join(d, "out")
        os.makedirs(out)
        env = dict(os.environ, OT2D_OUT=out, OT2D_NSTEPS=str(nsteps))
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([HARNESS], cwd=d, stdin=fin, env=env, check=True, stdout=subprocess.DEVNULL)
        norms = np.loadtxt(os.path.join(out, "norms.txt"), comments="#", ndmin=2)
        qfinal = np.fromfile(os.path.join(out, "qfinal_m1.bin"))
    deck_arr = {f"deck_{k}": np.asarray(float(v)) for k, v in deck.items()}
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), xv=g.xv, yv=g.yv, jtail1=np.asarray(g.jtail1),
                        nsteps=np.asarray(nsteps), norms=norms, qfinal=qfinal, **deck_arr)
    print(name, "norm[-1]", norms[-1, 3])


# Driver fixture: the reference program itself (oracle/_ref/overturns2d,
# src/Turns_overset.f main) on a small SA deck; its fort.71 (norms), fort.11
# (forces) and fort.8 (restart solution) are what ot2d_run must reproduce.
PROGRAM = os.path.join(ROOT, "oracle", "_ref", "overturns2d")
DRIVER_CASE = dict(jmax=41, kmax=21, m=0.02, p=0.4, t=0.12, ds_wall=1e-4, nsteps=12, nrest=5, npnorm=4)


def driver_deck():
    c = DRIVER_CASE
    g = grid.naca4_cgrid(c["jmax"], c["kmax"], m=c["m"], p=c["p"], t=c["t"], ds_wall=c["ds_wall"])
    return g, dict(SA, nsteps=c["nsteps"], nrest=c["nrest"], npnorm=c["npnorm"])


def _program_outputs(d, names):
    """The named output files of a run in directory d, as bytes (absent: not
    stored)."""
    out = {}
    for n in names:
        path = os.path.join(d, n)
        if os.path.exists(path):
            out["file:" + n] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    return out


def run_driver_case():
    g, deck = driver_deck()
    with tempfile.TemporaryDirectory() as d:
        grid.write_case(d, g, **deck)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([PROGRAM], cwd=d, stdin=fin, check=True, stdout=subprocess.DEVNULL)
        norms = np.loadtxt(os.path.join(d, "fort.71"), ndmin=2)
        forces = np.loadtxt(os.path.join(d, "fort.11"), ndmin=2)
        q8 = open(os.path.join(d, "fort.8"), "rb").read()
        files = _program_outputs(d, ["fort.9", "cp_TS1.dat", "SA.dat"])
    np.savez_compressed(os.path.join(HERE, "driver_sa_41x21.npz"), norms=norms, forces=forces,
                        fort8=np.frombuffer(q8, dtype=np.uint8), **files)
    print("driver_sa_41x21 norms", norms[-1])


# The reference program on a time-spectral deck (N = 3 instances of a
# pitching SA case) and on an overset deck (SST + WENO5 near-body C-grid in a
# Cartesian background): every output file the run leaves.
DRIVER_TS = dict(jmax=41, kmax=21, ds_wall=2e-4, nsteps=12, nrest=5, npnorm=4,
                 deck=dict(SA, fsmach=0.3, alfa=1.0, rey=4.0e6, dt=5.0, timespectral=True, nspec=3,
                           motiontype=1, rf=0.2, angmax=3.0))
DRIVER_OVERSET = dict(near=(97, 25), back=(61, 61), nsteps=8, nrest=5, npnorm=2,
                      deck=dict(SA, iturb=2, ilim=4, fsmach=0.3, alfa=6.0, rey=3.0e6, dt=5.0))


def driver_ts_deck():
    c = DRIVER_TS
    g = grid.naca4_cgrid(c["jmax"], c["kmax"], ds_wall=c["ds_wall"])
    return g, dict(c["deck"], nsteps=c["nsteps"], nrest=c["nrest"], npnorm=c["npnorm"])


def driver_overset_deck():
    c = DRIVER_OVERSET
    meshes = grid.airfoil_overset(near=c["near"], back=c["back"], ds_wall=1e-4)
    return meshes, dict(c["deck"], nsteps=c["nsteps"], nrest=c["nrest"], npnorm=c["npnorm"])


def run_driver_ts_case():
    g, deck = driver_ts_deck()
    n = deck["nspec"]
    with tempfile.TemporaryDirectory() as d:
        grid.write_case(d, g, **deck)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([PROGRAM], cwd=d, stdin=fin, check=True, stdout=subprocess.DEVNULL)
        names = ["fort.71", "fort.11"] + [f"{s}{i}{e}" for i in range(1, n + 1)
                                          for s, e in (("fort_TS", ".8"), ("fort_TS", ".9"), ("cp_TS", ".dat"),
                                                       ("SA_TS", ".dat"))]
        files = _program_outputs(d, names)
    np.savez_compressed(os.path.join(HERE, "driver_ts3_sa_41x21.npz"), **files)
    print("driver_ts3_sa_41x21", sorted(files))


def run_driver_overset_case():
    meshes, deck = driver_overset_deck()
    with tempfile.TemporaryDirectory() as d:
        grid.write_overset_case(d, meshes, **deck)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([PROGRAM], cwd=d, stdin=fin, check=True, stdout=subprocess.DEVNULL,
                           env=dict(os.environ, **ZERO_MALLOC))
        files = _program_outputs(d, ["fort.71", "fort.72", "fort.11", "fort.8", "fort.9", "cp_TS1.dat",
                                     "komega.dat"])
    np.savez_compressed(os.path.join(HERE, "driver_overset_sst_97x25_61x61.npz"), **files)
    print("driver_overset_sst_97x25_61x61", sorted(files))


# The reference program on a moving overset deck: the near-body SA C-grid
# pitches (iunst = 1, motiontype = 1) inside the fixed Cartesian background,
# and connect2d runs again after every step's motion
# (src/Turns_overset.f:411-465).
DRIVER_OVERSET_PITCH = dict(near=(97, 25), back=(61, 61), nsteps=6, nrest=4, npnorm=1,
                            deck=dict(SA, fsmach=0.3, alfa=2.0, rey=3.0e6, iunst=1, motiontype=1, rf=1.0,
                                      angmax=4.0, dt=0.1, iconstant_cfl=0))


def driver_overset_pitch_deck(static_conn=False):
    c = DRIVER_OVERSET_PITCH
    meshes = grid.airfoil_overset(near=c["near"], back=c["back"], ds_wall=1e-4)
    extra = dict(static_conn=True) if static_conn else {}
    return meshes, dict(c["deck"], nsteps=c["nsteps"], nrest=c["nrest"], npnorm=c["npnorm"], **extra)


# static_conn = .true.: the reference keeps the first connectivity while the
# near-body mesh moves (src/Turns_overset.f:413)
OVERSET_PITCH_FIXTURES = {"driver_pitch_overset_sa_97x25_61x61": False,
                          "driver_pitch_overset_static_sa_97x25_61x61": True}


def run_driver_overset_pitch_case(name):
    meshes, deck = driver_overset_pitch_deck(OVERSET_PITCH_FIXTURES[name])
    with tempfile.TemporaryDirectory() as d:
        grid.write_overset_case(d, meshes, **deck)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([PROGRAM], cwd=d, stdin=fin, check=True, stdout=subprocess.DEVNULL,
                           env=dict(os.environ, **ZERO_MALLOC))
        files = _program_outputs(d, ["fort.71", "fort.72", "fort.11", "fort.8", "fort.9", "cp_TS1.dat", "SA.dat"])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **files)
    print(name, sorted(files))


# Time-accurate decks (fixed grids, iunst = 0): Newton sub-iterations with
# the time term of src/time.f:174-209 and BDF2/BDF3 levels (stqol,
# src/time.f:109-172); the reference program's output files.
# name: (jmax, kmax, ds_wall, deck, nsteps, nrest, npnorm, zero-filled allocations)
DRIVER_TIME = {
    "driver_newton_sa_41x21": (41, 21, 1e-4, dict(SA, itnmax=3, ntac=2, dt=0.5, iconstant_cfl=0), 8, 5, 2, False),
    "driver_dual_sst_41x21": (41, 21, 1e-4, dict(SA, iturb=2, itnmax=4, ntac=3, idual=1, dt=0.2,
                                                  dtpseudo=[2.0, 1.5, 1.0, 0.5]), 7, 4, 1, True),
    "driver_newton_euler_33x17": (33, 17, 1e-3, dict(EULER, itnmax=2, ntac=1, dt=1.0), 6, 4, 2, False),
    # pitching grids (iunst = 1, motiontype = 1, src/motion.f:546-632)
    "driver_pitch_sa_41x21": (41, 21, 1e-4, dict(SA, iunst=1, motiontype=1, rf=0.2, angmax=2.0, dt=0.05,
                                                  iconstant_cfl=0), 8, 5, 2, False),
    "driver_pitch_euler_33x17": (33, 17, 1e-3, dict(EULER, iunst=1, motiontype=1, rf=0.3, angmax=3.0, dt=0.1,
                                                     ntac=2, itnmax=3, iconstant_cfl=0), 7, 4, 1, False),
    "driver_pitch_sst_41x21": (41, 21, 1e-4, dict(SA, iturb=2, iunst=1, motiontype=1, rf=0.25, angmax=1.5,
                                                   dt=0.05, ntac=2, itnmax=2, idual=1, dtpseudo=[1.0, 0.5]),
                               6, 4, 2, True),
}


def driver_time_deck(name):
    jm, km, ds, deck, nsteps, nrest, npnorm, _ = DRIVER_TIME[name]
    g = grid.naca4_cgrid(jm, km, m=0.02, p=0.4, t=0.12, ds_wall=ds)
    return g, dict(deck, nsteps=nsteps, nrest=nrest, npnorm=npnorm)


def run_driver_time_case(name):
    g, deck = driver_time_deck(name)
    zero = DRIVER_TIME[name][-1]
    with tempfile.TemporaryDirectory() as d:
        grid.write_case(d, g, **deck)
        env = dict(os.environ, **ZERO_MALLOC) if zero else dict(os.environ)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([PROGRAM], cwd=d, stdin=fin, check=True, stdout=subprocess.DEVNULL, env=env)
        files = _program_outputs(d, ["fort.71", "fort.11", "fort.8", "fort.9", "cp_TS1.dat", "SA.dat", "komega.dat"])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **files)
    print(name, sorted(files))


# Overset fixtures (num_grids = 2): connectivity by the reference's connect2d
# and, for the small pair, a trajectory with do_interpolations after every
# iteration.  name: (near-body jmax, kmax, background jmax, kmax, deck, nsteps)
SST_WENO = dict(SA, iturb=2, ilim=4, fsmach=0.3, alfa=6.0, rey=3.0e6, dt=5.0)
OVERSET_CASES = {
    "overset_97x25_61x61": (97, 25, 61, 61, SST_WENO, 10),
    "ihc_385x97_401x401": (385, 97, 401, 401, SST_WENO, 0),
}


def run_overset_case(name, nj, nk, bj, bk, deck, nsteps):
    meshes = grid.airfoil_overset(near=(nj, nk), back=(bj, bk), ds_wall=1e-4)
    with tempfile.TemporaryDirectory() as d:
        grid.write_overset_case(d, meshes, nsteps=max(nsteps, 1), nrest=10**6, npnorm=1, **deck)
        out = os.path.join(d, "out")
        os.makedirs(out)
        env = dict(os.environ, OT2D_OUT=out, OT2D_NSTEPS=str(nsteps), **ZERO_MALLOC)
        with open(os.path.join(d, "inputs")) as fin:
            subprocess.run([HARNESS], cwd=d, stdin=fin, env=env, check=True, stdout=subprocess.DEVNULL)
        arrays = {}
        for im in (1, 2):
            m = meshes[im - 1]
            if nsteps:
                arrays[f"xv{im}"], arrays[f"yv{im}"] = m.xv, m.yv
            else:
                # large pair: the grid is regenerated by grid.airfoil_overset,
                # pinned by the digest of its vertices
                h = hashlib.sha256(np.ascontiguousarray(m.xv).tobytes() + np.ascontiguousarray(m.yv).tobytes())
                arrays[f"grid_sha{im}"] = np.asarray(h.hexdigest())
            arrays[f"bc{im}"] = np.array([[b.ibtyp, b.ibdir, b.jbcs, b.jbce, b.kbcs, b.kbce] for b in m.bc])
            arrays[f"ihc{im}"] = np.array([[b.ibtyp, b.ibdir, b.jbcs, b.jbce, b.kbcs, b.kbce] for b in m.ihc],
                                          dtype=np.int64).reshape(-1, 6)
            arrays[f"counts{im}"] = np.loadtxt(os.path.join(out, f"conn_s1_m{im}.txt"), dtype=np.int64)
            arrays[f"iblank{im}"] = np.fromfile(os.path.join(out, f"iblank_m{im}.bin"), dtype=np.int32)
            for key, dt in (("imesh", np.int32), ("ibc", np.int32), ("idonor", np.int32), ("frac", np.float64)):
                arrays[f"{key}{im}"] = np.fromfile(os.path.join(out, f"{key}_s1_m{im}.bin"), dtype=dt)
            if nsteps:
                arrays[f"qfinal{im}"] = np.fromfile(os.path.join(out, f"qfinal_m{im}.bin"))
                arrays[f"turmufinal{im}"] = np.fromfile(os.path.join(out, f"turmufinal_m{im}.bin"))
        if nsteps:
            arrays["norms"] = np.loadtxt(os.path.join(out, "norms.txt"), comments="#", ndmin=2)
    deck_arr = {f"deck_{k}": np.asarray(float(v)) for k, v in deck.items()}
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), nsteps=np.asarray(nsteps), jtail1=np.asarray(meshes[0].jtail1),
                        **arrays, **deck_arr)
    print(name, "fringes", arrays["counts1"][0], arrays["counts2"][0])


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: run `make -C oracle ref` first")
    only = set(sys.argv[1:])   # fixture names to (re)write; none: all
    for name, spec in CASES.items():
        if not only or name in only:
            run_case(name, *spec)
    for name, spec in TS_CASES.items():
        if not only or name in only:
            run_ts_case(name, *spec)
    if not only or "driver_sa_41x21" in only:
        run_driver_case()
    if not only or "driver_ts3_sa_41x21" in only:
        run_driver_ts_case()
    if not only or "driver_overset_sst_97x25_61x61" in only:
        run_driver_overset_case()
    for name in DRIVER_TIME:
        if not only or name in only:
            run_driver_time_case(name)
    for name in OVERSET_PITCH_FIXTURES:
        if not only or name in only:
            run_driver_overset_pitch_case(name)
    for name, spec in OVERSET_CASES.items():
        if not only or name in only:
            run_overset_case(name, *spec)
