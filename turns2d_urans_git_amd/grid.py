"""C-grid generator and input-deck writer for OVERTURNS-2D style cases.

The reference solver reads a vertex grid from ``fort.1`` (``src/io.f:343-375``,
dimensions read at ``src/setup.f:1253-1265``), a namelist ``&inputs`` deck
from stdin (``src/modules.f:197-208``, read at ``src/setup.f:908``) and the
``&bcinp`` boundary table from ``bc.inp`` (``src/setup.f:899``, ``911-931``).
It ships no grids or decks, so this module makes both:

* ``naca4_cgrid`` builds an algebraic C-grid around a NACA 4-digit airfoil with
  a wake cut on ``k = 1`` (vertex ``j`` and ``jmax-j+1`` coincide for
  ``j <= jtail1``), geometric wall-normal stretching from a given first
  spacing, and a C-shaped outer boundary at ``R`` chords.  ``j`` runs from the
  lower wake outflow, under the airfoil, round the leading edge and back along
  the upper wake, so the ``(j, k)`` frame is right-handed and every cell area
  is positive (``src/setup.f:489-500``).  The trailing edge is closed (the
  -0.1036 x^4 coefficient) because the wake cut needs coincident upper and
  lower trailing-edge vertices.
* ``cartesian_grid`` builds a uniform background mesh (overset background).
* ``write_fort1`` / ``read_fort1`` use Fortran sequential unformatted records
  (4-byte length markers) exactly as the reference reads them.
* ``cgrid_bc`` returns the boundary table the reference expects for this
  topology: wall (5), wake cut (51), characteristic far field (47) on the
  three outer sides.
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import Dict, List, Optional, Sequence

import numpy as np


# --------------------------------------------------------------------------
# airfoil geometry
# --------------------------------------------------------------------------
def naca4_surface(m: float, p: float, t: float, xc: np.ndarray):
    """Upper and lower NACA 4-digit surface at chordwise stations ``xc``.

    Thickness ``y_t = 5t(0.2969 sqrt(x) - 0.1260 x - 0.3516 x^2 + 0.2843 x^3
    - 0.1036 x^4)`` (closed trailing edge); camber ``m, p`` per the standard
    4-digit definition, thickness applied normal to the camber line.
    """
    xc = np.asarray(xc, dtype=np.float64)
    yt = 5.0 * t * (0.2969 * np.sqrt(np.maximum(xc, 0.0)) - 0.1260 * xc
                    - 0.3516 * xc**2 + 0.2843 * xc**3 - 0.1036 * xc**4)
    yc = np.zeros_like(xc)
    dyc = np.zeros_like(xc)
    if m > 0.0 and p > 0.0:
        f = xc < p
        yc[f] = m / p**2 * (2 * p * xc[f] - xc[f]**2)
        dyc[f] = 2 * m / p**2 * (p - xc[f])
        g = ~f
        yc[g] = m / (1 - p)**2 * ((1 - 2 * p) + 2 * p * xc[g] - xc[g]**2)
        dyc[g] = 2 * m / (1 - p)**2 * (p - xc[g])
    th = np.arctan(dyc)
    xu = xc - yt * np.sin(th)
    yu = yc + yt * np.cos(th)
    xl = xc + yt * np.sin(th)
    yl = yc - yt * np.cos(th)
    return xu, yu, xl, yl


def geometric_distribution(ds1: float, length: float, n: int) -> np.ndarray:
    """``n`` points from 0 to ``length`` with first spacing ``ds1`` and a
    constant growth ratio (solved by bisection).  Falls back to uniform
    spacing if ``ds1 * (n-1) >= length``."""
    if n < 2:
        return np.zeros(max(n, 0))
    nint = n - 1
    if ds1 * nint >= length:
        return np.linspace(0.0, length, n)

    def total(r):
        return ds1 * (r**nint - 1.0) / (r - 1.0)

    lo, hi = 1.0 + 1e-12, 2.0
    while total(hi) < length:
        hi *= 1.5
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) < length:
            lo = mid
        else:
            hi = mid
    r = 0.5 * (lo + hi)
    d = np.concatenate([[0.0], np.cumsum(ds1 * r**np.arange(nint))])
    d *= length / d[-1]
    return d


def growth_ratio(ds1: float, length: float, n: int) -> float:
    d = geometric_distribution(ds1, length, n)
    return float((d[2] - d[1]) / (d[1] - d[0])) if n > 2 else 1.0


@dataclasses.dataclass
class CGrid:
    xv: np.ndarray          # (jmax, kmax) float64, Fortran order
    yv: np.ndarray
    jtail1: int             # vertex index (1-based) of the lower trailing edge
    jmax: int
    kmax: int

    @property
    def jtail2(self) -> int:
        return self.jmax - self.jtail1 + 1


def naca4_cgrid(jmax: int = 257, kmax: int = 65, m: float = 0.0, p: float = 0.0,
                t: float = 0.12, ds_wall: float = 1e-4, radius: float = 20.0,
                jtail1: Optional[int] = None, wake_frac: float = 0.125,
                wake_aspect: Optional[float] = None, te_cluster: bool = True) -> CGrid:
    """Algebraic C-grid around a NACA 4-digit airfoil (see module docstring).

    ``jtail1`` defaults to about ``wake_frac * jmax`` wake points per side.
    Grid lines run straight from each inner point (airfoil or wake cut) to its
    outer-boundary point with geometric spacing from ``ds_wall``; straight
    lines between monotonically ordered inner and outer points never cross.
    With ``wake_aspect`` the first spacing off the wake cut grows with the
    streamwise spacing there, ``max(ds_wall, dx / wake_aspect)``, so cells at
    the downstream boundary do not reach aspect ratios of 1e6 on wall-resolved
    grids.  ``te_cluster=False`` clusters surface points at the leading edge
    only (x = 1 - cos(pi s / 2)), which keeps the streamwise spacing at the
    trailing edge near pi / (2 nside): the reference's implicit operator
    carries the viscous term of the wall-normal direction only
    (src/lhs.f:549-553), so tiny streamwise cells under a turbulent wake limit
    its stable CFL.
    """
    if jtail1 is None:
        jtail1 = max(3, int(round(wake_frac * jmax)) + 1)
    nsurf = jmax - 2 * (jtail1 - 1)          # surface vertices incl. both TE
    if nsurf < 5 or nsurf % 2 == 0:
        raise ValueError(f"jmax={jmax}, jtail1={jtail1} give {nsurf} surface points; need odd >= 5")
    nside = (nsurf - 1) // 2                 # intervals per side
    # cosine clustering at LE and TE
    if te_cluster:
        theta = np.linspace(0.0, np.pi, nside + 1)
        xc = 0.5 * (1.0 + np.cos(theta))     # 1 -> 0  (TE -> LE)
    else:
        sgrid = np.linspace(1.0, 0.0, nside + 1)
        xc = 1.0 - np.cos(0.5 * np.pi * sgrid)
        xc[0] = 1.0
    xu, yu, xl, yl = naca4_surface(m, p, t, xc)
    # lower surface TE->LE, then upper surface LE->TE
    sx = np.concatenate([xl, xu[::-1][1:]])
    sy = np.concatenate([yl, yu[::-1][1:]])
    sx[0] = sx[-1] = 1.0
    sy[0] = sy[-1] = 0.0
    # wake: from far downstream to the TE, spacing grows away from the TE
    ds_te = float(np.hypot(sx[1] - sx[0], sy[1] - sy[0]))
    wd = geometric_distribution(ds_te, radius, jtail1)   # 0 at TE .. radius
    wx_lower = 1.0 + wd[::-1]                # j = 1 .. jtail1 (x decreasing)
    x0 = np.concatenate([wx_lower, sx[1:-1], wx_lower[::-1]])
    y0 = np.concatenate([np.zeros(jtail1), sy[1:-1], np.zeros(jtail1)])
    assert x0.size == jmax
    # outer boundary
    x1 = np.empty(jmax)
    y1 = np.empty(jmax)
    x1[:jtail1] = wx_lower
    y1[:jtail1] = -radius
    x1[jmax - jtail1:] = wx_lower[::-1]
    y1[jmax - jtail1:] = radius
    # airfoil columns map to the semicircle centred on the TE
    js = np.arange(jtail1 - 1, jmax - jtail1 + 1)    # 0-based incl. both TE
    frac = (js - js[0]) / (js[-1] - js[0])
    psi = -0.5 * np.pi - np.pi * frac
    x1[js] = 1.0 + radius * np.cos(psi)
    y1[js] = radius * np.sin(psi)
    x1[jtail1 - 1] = 1.0
    y1[jtail1 - 1] = -radius
    x1[jmax - jtail1] = 1.0
    y1[jmax - jtail1] = radius

    xv = np.empty((jmax, kmax), order="F")
    yv = np.empty((jmax, kmax), order="F")
    ds = np.full(jmax, ds_wall)
    if wake_aspect:
        dxw = np.abs(np.diff(wx_lower))                 # lower wake, j = 1 .. jtail1
        dxj = np.concatenate([dxw[:1], np.maximum(dxw[:-1], dxw[1:]) if dxw.size > 1 else dxw, [0.0]])
        dsw = np.maximum(ds_wall, dxj / wake_aspect)
        ds[:jtail1] = dsw
        ds[jmax - jtail1:] = dsw[::-1]
    for j in range(jmax):
        L = float(np.hypot(x1[j] - x0[j], y1[j] - y0[j]))
        d = geometric_distribution(float(ds[j]), L, kmax) / L
        xv[j, :] = x0[j] + d * (x1[j] - x0[j])
        yv[j, :] = y0[j] + d * (y1[j] - y0[j])
    # wake cut: make mirrored vertices bit-identical
    for j in range(jtail1):
        jj = jmax - 1 - j
        xv[jj, 0] = xv[j, 0]
        yv[jj, 0] = yv[j, 0]
    return CGrid(xv=xv, yv=yv, jtail1=jtail1, jmax=jmax, kmax=kmax)


def cartesian_grid(jmax: int, kmax: int, xlo: float, xhi: float, ylo: float, yhi: float):
    x = np.linspace(xlo, xhi, jmax)
    y = np.linspace(ylo, yhi, kmax)
    xv = np.asfortranarray(np.repeat(x[:, None], kmax, axis=1))
    yv = np.asfortranarray(np.repeat(y[None, :], jmax, axis=0))
    return xv, yv


def cell_areas(xv: np.ndarray, yv: np.ndarray) -> np.ndarray:
    """Cell areas with the reference's diagonal cross product
    (``src/setup.f:489-500``); positive for a valid grid."""
    dx1 = xv[1:, 1:] - xv[:-1, :-1]
    dy1 = yv[1:, 1:] - yv[:-1, :-1]
    dx2 = xv[:-1, 1:] - xv[1:, :-1]
    dy2 = yv[:-1, 1:] - yv[1:, :-1]
    return 0.5 * (dx1 * dy2 - dx2 * dy1)


# --------------------------------------------------------------------------
# Fortran unformatted sequential records
# --------------------------------------------------------------------------
def _write_record(f, payload: bytes):
    n = len(payload)
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def _read_record(f) -> bytes:
    h = f.read(4)
    if len(h) < 4:
        raise EOFError
    (n,) = struct.unpack("<i", h)
    data = f.read(n)
    (n2,) = struct.unpack("<i", f.read(4))
    if n2 != n:
        raise ValueError("corrupt Fortran record")
    return data


def write_fort1(path: str, grids: Sequence[tuple]):
    """Write ``fort.1``: ``[nmesh]`` (only when several meshes), the
    ``(jmax,kmax)`` list, then one record of ``xv, yv`` per mesh
    (``src/setup.f:1253-1255``, ``src/io.f:359-366``)."""
    with open(path, "wb") as f:
        if len(grids) > 1:
            _write_record(f, struct.pack("<i", len(grids)))
        dims = b"".join(struct.pack("<ii", g[0].shape[0], g[0].shape[1]) for g in grids)
        _write_record(f, dims)
        for xv, yv in grids:
            _write_record(f, np.asfortranarray(xv, dtype=np.float64).tobytes(order="F")
                          + np.asfortranarray(yv, dtype=np.float64).tobytes(order="F"))


def read_fort1(path: str, num_grids: int = 1):
    with open(path, "rb") as f:
        if num_grids > 1:
            (nmesh,) = struct.unpack("<i", _read_record(f))
        else:
            nmesh = 1
        dims = np.frombuffer(_read_record(f), dtype="<i4").reshape(nmesh, 2)
        out = []
        for jm, km in dims:
            buf = np.frombuffer(_read_record(f), dtype="<f8")
            n = jm * km
            out.append((buf[:n].reshape((jm, km), order="F").copy(),
                        buf[n:2 * n].reshape((jm, km), order="F").copy()))
    return out


def read_qfile(path: str, jd: int, kd: int, nturb: int = 0, num_grids: int = 1,
               dims=None):
    """Read a reference restart/solution file (``fort.8``) as written by
    ``astore`` (``src/io.f:312-320``): header records then ``q(jd,kd,1:4)``
    (conserved, not area-scaled) and the turbulence variables."""
    with open(path, "rb") as f:
        if num_grids > 1:
            _read_record(f)
        _read_record(f)                       # dims
        hdr = np.frombuffer(_read_record(f), dtype="<f8")
        q = np.frombuffer(_read_record(f), dtype="<f8").reshape((jd, kd, 4), order="F").copy()
        turb = None
        if nturb:
            turb = np.frombuffer(_read_record(f), dtype="<f8").reshape((jd, kd, nturb), order="F").copy()
    return hdr, q, turb


# --------------------------------------------------------------------------
# decks
# --------------------------------------------------------------------------
@dataclasses.dataclass
class BCSeg:
    ibtyp: int
    ibdir: int
    jbcs: int
    jbce: int
    kbcs: int
    kbce: int


def cgrid_bc(jtail1: int) -> List[BCSeg]:
    """Boundary table for ``naca4_cgrid``: wall, wake cut, far field at
    ``k = kmax``, ``j = 1`` and ``j = jmax`` (negative indices count from the
    end, ``src/bc.f:25-28``)."""
    return [
        BCSeg(5, 2, jtail1, -jtail1, 1, 1),
        BCSeg(51, 2, 1, jtail1, 1, 1),
        BCSeg(47, -2, 1, -1, -1, -1),
        BCSeg(47, 1, 1, 1, 1, -1),
        BCSeg(47, -1, -1, -1, 1, -1),
    ]


def _fmt(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v))
    if isinstance(v, (list, tuple)):
        return ",".join(_fmt(x) for x in v)
    raise TypeError(v)


def write_inputs(path: str, **nml):
    """Write an ``&inputs`` namelist deck (names as in ``src/modules.f:197-208``)."""
    with open(path, "w") as f:
        f.write(" &inputs\n")
        for k, v in nml.items():
            f.write(f"  {k} = {_fmt(v)},\n")
        f.write(" /\n")


def write_bcinp(path: str, meshes: Sequence[Sequence[BCSeg]], irot_dir: int = 1):
    """Write ``bc.inp``: one ``&bcinp`` group per mesh (``src/setup.f:911-931``)."""
    with open(path, "w") as f:
        for segs in meshes:
            f.write(" &bcinp\n")
            f.write(f"  nbc = {len(segs)},\n")
            f.write(f"  irot_dir = {irot_dir},\n")
            for name in ("ibtyp", "ibdir", "jbcs", "jbce", "kbcs", "kbce"):
                vals = ",".join(str(getattr(s, name)) for s in segs)
                f.write(f"  {name} = {vals},\n")
            f.write(" /\n")


def write_case(dirpath: str, grid: CGrid, segs: Optional[List[BCSeg]] = None, **nml):
    """Write ``fort.1``, ``bc.inp`` and the ``inputs`` deck for one C-grid case."""
    os.makedirs(dirpath, exist_ok=True)
    write_fort1(os.path.join(dirpath, "fort.1"), [(grid.xv, grid.yv)])
    write_bcinp(os.path.join(dirpath, "bc.inp"), [segs or cgrid_bc(grid.jtail1)])
    base = dict(jmax=grid.jmax, kmax=grid.kmax, jtail1=grid.jtail1)
    base.update(nml)
    write_inputs(os.path.join(dirpath, "inputs"), **base)


# --------------------------------------------------------------------------
# overset decks
# --------------------------------------------------------------------------
@dataclasses.dataclass
class IHCSeg:
    """One row of an ``ihcbc.<n>`` file (``ihc/src/ihc.f90:147-208``).

    ``ibtyp``: 1 not a receiver, 2 blanked, 3 forced receiver, 4 never
    blanked, 5 wall points (vertex-count indexing), 6 = 1 and 4, 7 = 3 and 4.
    Ranges are cell indices of the mesh without halo (negative counts from
    the end)."""
    ibtyp: int
    ibdir: int
    jbcs: int
    jbce: int
    kbcs: int
    kbce: int


def write_ihcbc(path: str, segs: Sequence[IHCSeg]):
    """Write one ``&ihcbcinp`` namelist file."""
    with open(path, "w") as f:
        f.write(" &ihcbcinp\n")
        f.write(f"  nbc = {len(segs)},\n")
        if segs:
            for name in ("ibtyp", "ibdir", "jbcs", "jbce", "kbcs", "kbce"):
                vals = ",".join(str(getattr(sg, name)) for sg in segs)
                f.write(f"  {name} = {vals},\n")
        f.write(" /\n")


def box_bc() -> List[BCSeg]:
    """Characteristic far field (type 47) on the four sides of a Cartesian
    background mesh."""
    return [
        BCSeg(47, 2, 1, -1, 1, 1),
        BCSeg(47, -2, 1, -1, -1, -1),
        BCSeg(47, 1, 1, 1, 1, -1),
        BCSeg(47, -1, -1, -1, 1, -1),
    ]


def cgrid_ihc(jtail1: int, jmax: int, fringe_layers: int = 2) -> List[IHCSeg]:
    """IHC rows for a near-body C-grid: the airfoil wall points (cells on
    ``k = 1`` between the trailing-edge vertices) and forced receivers on the
    outermost ``fringe_layers`` cell layers."""
    return [
        IHCSeg(5, 2, jtail1, jmax - jtail1, 1, 1),
        IHCSeg(3, -2, 1, -1, -fringe_layers, -1),
    ]


@dataclasses.dataclass
class OversetMesh:
    xv: np.ndarray
    yv: np.ndarray
    bc: List[BCSeg]
    ihc: List[IHCSeg]
    jtail1: int = 1


def airfoil_overset(near=(385, 97), back=(401, 401), m=0.02, p=0.4, t=0.12, ds_wall=1e-5,
                    radius=1.5, half_width=10.0, fringe_layers=2, **kw) -> List[OversetMesh]:
    """Near-body C-grid around a NACA 4-digit airfoil (outer radius
    ``radius`` chords) inside a uniform Cartesian background over
    ``[-half_width, half_width]^2``."""
    g = naca4_cgrid(near[0], near[1], m=m, p=p, t=t, ds_wall=ds_wall, radius=radius, **kw)
    xb, yb = cartesian_grid(back[0], back[1], -half_width, half_width, -half_width, half_width)
    return [OversetMesh(g.xv, g.yv, cgrid_bc(g.jtail1), cgrid_ihc(g.jtail1, g.jmax, fringe_layers), g.jtail1),
            OversetMesh(xb, yb, box_bc(), [], 1)]


def write_overset_case(dirpath: str, meshes: Sequence[OversetMesh], **nml):
    """Write ``fort.1``, ``bc.inp``, ``ihcbc.<n>`` and ``inputs`` for an
    overset case (``num_grids`` meshes)."""
    os.makedirs(dirpath, exist_ok=True)
    write_fort1(os.path.join(dirpath, "fort.1"), [(mm.xv, mm.yv) for mm in meshes])
    write_bcinp(os.path.join(dirpath, "bc.inp"), [mm.bc for mm in meshes])
    for i, mm in enumerate(meshes):
        write_ihcbc(os.path.join(dirpath, f"ihcbc.{i}"), mm.ihc)
    base = dict(num_grids=len(meshes), jtail1=meshes[0].jtail1)
    base.update(nml)
    write_inputs(os.path.join(dirpath, "inputs"), **base)
