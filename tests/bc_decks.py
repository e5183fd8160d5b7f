"""Decks for the boundary-condition types and stop paths that the airfoil
C-grids never reach: extrapolation (type 10, bcextp, src/bc.f:1962-2026),
symmetry (11, bcsym, src/bc.f:2096-2157), periodic (22, bcperiodic,
src/bc.f:180-267) and free stream (46, bcout_inf, src/bc.f:360-407), with
their residual branches in rbc (src/bc.f:88-177) and the turbulence models'
boundary branches (src/samodel2d.f:864-890, src/komegasst.f:943-970); and
the two ways the reference stops: monitor's negative speed of sound
(src/step.f:332-392, src/Turns_overset.f:495-497) and the residual norm out
of bounds (stop 'norm', src/step.f:105-110).

All run on a 41 x 21 Cartesian box from a perturbed state (the free stream
itself is steady under these boundary conditions).  Shared by the live
reference tests (test_oracle_ref.py) and the GPU tests (test_gpu_bc_stop.py).
"""
from __future__ import annotations

import numpy as np

from turns2d_urans_git_amd import grid
from turns2d_urans_git_amd.grid import BCSeg

GAMMA = 1.4

EULER = dict(fsmach=0.4, alfa=0.0, invisc=1, lamin=1, irhsy=-3, ilim=1, ilhs=1, iconstant_cfl=1, dt=10.0,
             epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0)
SA = dict(fsmach=0.4, alfa=0.0, rey=1.0e5, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3, ilim=1, ilhs=1,
          iconstant_cfl=1, dt=5.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0)
SST = dict(SA, iturb=2)

# k=1 symmetry, k=kmax extrapolation, j=1 free stream, j=jmax extrapolation
EXTP_SYM_INF = [BCSeg(11, 2, 1, -1, 1, 1), BCSeg(10, -2, 1, -1, -1, -1), BCSeg(46, 1, 1, 1, 1, -1),
                BCSeg(10, -1, -1, -1, 1, -1)]
# the other directions of each type
EXTP_SYM_OTHER = [BCSeg(10, 2, 1, -1, 1, 1), BCSeg(11, -2, 1, -1, -1, -1), BCSeg(11, 1, 1, 1, 1, -1),
                  BCSeg(46, -1, -1, -1, 1, -1)]
# periodic in j (between the first and last interior columns' images)
PERIODIC_J = [BCSeg(22, 1, 1, 1, 1, -1), BCSeg(22, -1, -1, -1, 1, -1), BCSeg(11, 2, 1, -1, 1, 1),
              BCSeg(46, -2, 1, -1, -1, -1)]
# periodic in k
PERIODIC_K = [BCSeg(22, 2, 1, -1, 1, 1), BCSeg(22, -2, 1, -1, -1, -1), BCSeg(46, 1, 1, 1, 1, -1),
              BCSeg(10, -1, -1, -1, 1, -1)]
# no inflow: extrapolation on three sides, symmetry on k=1 (for the stops)
CLOSED = [BCSeg(11, 2, 1, -1, 1, 1), BCSeg(10, -2, 1, -1, -1, -1), BCSeg(10, 1, 1, 1, 1, -1),
          BCSeg(10, -1, -1, -1, 1, -1)]

BC_DECKS = {
    "euler_extp_sym_inf": (EULER, EXTP_SYM_INF),
    "euler_extp_sym_other": (EULER, EXTP_SYM_OTHER),
    "euler_periodic_j": (EULER, PERIODIC_J),
    "euler_periodic_k": (EULER, PERIODIC_K),
    "sa_extp_sym_inf": (SA, EXTP_SYM_INF),
    "sa_periodic_j": (SA, PERIODIC_J),
    "sst_extp_sym_inf": (SST, EXTP_SYM_INF),
    "sst_periodic_k": (SST, PERIODIC_K),
}

# monitor stops: a quiescent gas whose speed of sound squared lies near the
# 0.002 floor (a^2, velocity amplitude) -> the iteration the reference stops at
MSTOP_CASES = {"first_step": (0.0019, 0.005, 1), "second_step": (0.004, 0.05, 2)}


def box(jmax: int = 41, kmax: int = 21) -> grid.CGrid:
    xv, yv = grid.cartesian_grid(jmax, kmax, -1.0, 1.0, -0.5, 0.5)
    return grid.CGrid(xv=np.asfortranarray(xv), yv=np.asfortranarray(yv), jtail1=1, jmax=jmax, kmax=kmax)


def perturbed(q: np.ndarray, amp: float = 0.05) -> np.ndarray:
    """Smooth perturbation of a free-stream state q(jd,kd,1,nq): conserved
    variables scaled by 1 +- amp, a cross-flow added, turbulence variables
    scaled by 1 +- amp/2 (q(nq) = 1/area is left alone)."""
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    jd, kd, _, nq = q.shape
    J, K = np.meshgrid(np.arange(jd), np.arange(kd), indexing="ij")
    for n in range(4):
        q[:, :, 0, n] *= 1 + amp * np.sin(0.37 * J + 0.5 * n) * np.cos(0.53 * K - 0.2 * n)
    q[:, :, 0, 2] += 0.03 * q[:, :, 0, 0] * np.sin(0.2 * J)
    for n in range(4, nq - 1):
        q[:, :, 0, n] *= 1 + 0.5 * amp * np.cos(0.29 * J + n) * np.sin(0.41 * K + 0.3)
    return q


def quiescent(q: np.ndarray, asq: float, amp: float) -> np.ndarray:
    """A gas at rest but for a velocity field of amplitude amp, with speed of
    sound squared asq (density 1 +- 2%), in the area scaling of q."""
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    jd, kd = q.shape[:2]
    J, K = np.meshgrid(np.arange(jd), np.arange(kd), indexing="ij")
    area = 1.0 / q[:, :, 0, -1]
    r = 1.0 + 0.02 * np.sin(0.3 * J) * np.cos(0.4 * K)
    u = amp * np.sin(0.35 * J + 0.2) * np.cos(0.25 * K)
    v = amp * np.cos(0.3 * J) * np.sin(0.45 * K + 0.1)
    p = asq * r / GAMMA
    e = p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)
    q[:, :, 0, 0] = r * area
    q[:, :, 0, 1] = r * u * area
    q[:, :, 0, 2] = r * v * area
    q[:, :, 0, 3] = e * area
    return q


def blown_up(q: np.ndarray, factor: float = 1.0e5) -> np.ndarray:
    """perturbed(q) with the conserved variables of a 10 x 6 block scaled by
    `factor`: the first residual's norm exceeds 1000."""
    q = perturbed(q)
    q[15:25, 8:14, 0, :4] *= factor
    return q
