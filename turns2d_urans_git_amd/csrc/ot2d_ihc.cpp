// Implicit hole cutting and donor search for overset meshes (connect2d and
// its helpers, ihc/src/ihc.f90:252-1484, with the octree of
// ihc/src/octree.f90:43-253).
//
// Runs once per static overset configuration, on the host: the search walks
// every receiver candidate in a fixed serpentine order and warm-starts each
// donor walk from the previous hit, so its outcome (which cells are cut,
// which donor each receiver gets, and the order of the lists) depends on
// that order.  The results are the integer lists and quadratic-Lagrange
// fractions the per-iteration interpolation kernels consume on the device.
//
// Index conventions follow the reference module: vertices (j,k) with
// 1 <= j <= jmax, cells 0..jmax x 0..kmax including the halo ring (cell 0 is
// the halo), interior cells 1..jmaxc = jmax-1.  Arithmetic is operand for
// operand the reference's (built with -ffp-contract=off).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "ot2d_ihc.h"

namespace ot2d {
namespace ihc {

namespace {

constexpr double kOverlap = 0.95;   // overlap_factor (ihc/src/ihc.f90:135)
constexpr int kLvlMin = 2, kLvlMax = 5, kMaxLev = 10;

// ---------------------------------------------------------------- octree
// Morton index of a point in [0,1)^2 at level l, the reference's bit order:
// bit 2i holds bit i of the y cell, bit 2i+1 bit i of the x cell
// (localboxindex/interleave, octree.f90:187-242), then the global box number
// (index_ltog, octree.f90:206-216): levels are numbered consecutively from 1.
int64_t ltog(int64_t n, int l) {
  int64_t m = n + 1;
  for (int i = 0; i < l; ++i) m += int64_t(1) << (2 * i);
  return m;
}

int64_t box_index(const double xbar[2], int l) {
  int64_t c[2];
  for (int d = 0; d < 2; ++d) c[d] = (int64_t)((double)(int64_t(1) << l) * xbar[d]);
  int64_t out = 0;
  int count = 0;
  while (c[0] != 0 || c[1] != 0) {
    for (int d = 1; d >= 0; --d) {
      int64_t bit = c[d] & 1;
      out += bit << count;
      ++count;
      c[d] >>= 1;
    }
  }
  return ltog(out, l);
}

struct Octree {
  double xs[2], scale[2];
  int64_t nbox;
  std::vector<uint8_t> empty;
  std::vector<int> lvlpoint;
  std::vector<int> istart[2], iend[2];
};

// make_octree for one mesh (octree.f90:43-172)
void make_octree(const MeshIn& m, Octree& t) {
  t.nbox = 0;
  for (int l = 0; l <= kLvlMax; ++l) t.nbox += int64_t(1) << (2 * l);
  t.empty.assign(t.nbox + 1, 1);
  t.lvlpoint.assign(t.nbox + 1, 0);
  for (int d = 0; d < 2; ++d) {
    t.istart[d].assign(t.nbox + 1, 1);
    t.iend[d].assign(t.nbox + 1, 1);
  }
  double lo[2] = {m.xv[0], m.yv[0]}, hi[2] = {m.xv[0], m.yv[0]};
  for (int k = 1; k <= m.kmax; ++k)
    for (int j = 1; j <= m.jmax; ++j) {
      double x = m.XV(j, k), y = m.YV(j, k);
      lo[0] = x < lo[0] ? x : lo[0]; hi[0] = x > hi[0] ? x : hi[0];
      lo[1] = y < lo[1] ? y : lo[1]; hi[1] = y > hi[1] ? y : hi[1];
    }
  double ex = (hi[0] - lo[0]) * 0.0001, ey = (hi[1] - lo[1]) * 0.0001;
  double xs0 = lo[0] - ex, xs1 = lo[1] - ey, xe0 = hi[0] + ex, xe1 = hi[1] + ey;
  t.xs[0] = xs0; t.xs[1] = xs1;
  t.scale[0] = xe0 - xs0; t.scale[1] = xe1 - xs1;
  // first and last vertex (in the finest-level order) of every leaf box; the
  // running bounds are those of the most recently opened box, as in the
  // reference (octree.f90:85-112)
  int64_t nsbox = 0, nebox = 0;
  for (int j = 1; j <= m.jmax; ++j)
    for (int k = 1; k <= m.kmax; ++k) {
      double xb[2] = {(m.XV(j, k) - t.xs[0]) / t.scale[0], (m.YV(j, k) - t.xs[1]) / t.scale[1]};
      int64_t nout = box_index(xb, kLvlMax);
      if (t.empty[nout]) {
        t.istart[0][nout] = j; t.istart[1][nout] = k;
        nsbox = box_index(xb, kMaxLev);
        t.iend[0][nout] = j; t.iend[1][nout] = k;
        nebox = nsbox;
        t.empty[nout] = 0;
      } else {
        int64_t nm = box_index(xb, kMaxLev);
        if (nm < nsbox) {
          t.istart[0][nout] = j; t.istart[1][nout] = k;
          nsbox = nm;
        } else if (nm > nebox) {
          t.iend[0][nout] = j; t.iend[1][nout] = k;
          nebox = nm;
        }
      }
    }
  // parents inherit the first start and the last end of their children
  for (int l = kLvlMax; l >= 1; --l) {
    int64_t maxbox = int64_t(1) << (2 * l);
    for (int64_t b = 0; b < maxbox; ++b) {
      int64_t nout = ltog(b, l), npout = ltog(b >> 2, l - 1);
      if (!t.empty[nout]) {
        if (t.empty[npout]) { t.istart[0][npout] = t.istart[0][nout]; t.istart[1][npout] = t.istart[1][nout]; }
        t.empty[npout] = 0;
        t.iend[0][npout] = t.iend[0][nout];
        t.iend[1][npout] = t.iend[1][nout];
      }
    }
  }
  // level of the nearest non-empty ancestor of every leaf (down to lvlmin)
  int64_t maxbox = int64_t(1) << (2 * kLvlMax);
  for (int64_t b = 0; b < maxbox; ++b) {
    int64_t nout = ltog(b, kLvlMax);
    t.lvlpoint[nout] = kLvlMax;
    if (t.empty[nout]) {
      int64_t pb = b;
      for (int l = kLvlMax - 1; l >= kLvlMin; --l) {
        pb >>= 2;
        if (!t.empty[ltog(pb, l)]) { t.lvlpoint[nout] = l; break; }
      }
    }
  }
}

// ------------------------------------------------------------- finddonor
// Walk from cell (js,ks) of a mesh's vertex grid towards the cell holding xp
// (ihc/src/ihc.f90:1118-1248).  Found: js,ks > 0.  Left the grid: the last
// cell, negated.  Stopped after maxsearch steps: js < 0, ks > 0.
void finddonor(const double xp[2], const MeshIn& m, int& js, int& ks, int maxsearch) {
  static const int mo[4] = {1, 1, 2, 3}, ma[4] = {3, 2, 4, 4}, mb[4] = {2, 3, 1, 1};
  const int jmax = m.jmax, kmax = m.kmax;
  int j = js < jmax - 1 ? js : jmax - 1;
  int k = ks < kmax - 1 ? ks : kmax - 1;
  // starts are cells >= 1 on every path the callers take; kept in range so
  // the corner reads below stay inside the mesh
  j = j < 1 ? 1 : j;
  k = k < 1 ? 1 : k;
  bool inside = false, outside = false, notoutside = false;
  int nsearch = 0;
  if (maxsearch == 0) maxsearch = 60;
  int movejp = 0, movekp = 0, movej = 0, movek = 0;
  // the reference's cycle-detection pair starts uninitialised; any value
  // outside the grid behaves the same
  int i1 = -(1 << 30), i2 = -(1 << 30);
  while (!inside && !outside) {
    inside = true;
    const int jj = j, kk = k;
    double c[5][2];
    c[1][0] = m.XV(jj, kk);         c[1][1] = m.YV(jj, kk);
    c[2][0] = m.XV(jj + 1, kk);     c[2][1] = m.YV(jj + 1, kk);
    c[3][0] = m.XV(jj, kk + 1);     c[3][1] = m.YV(jj, kk + 1);
    c[4][0] = m.XV(jj + 1, kk + 1); c[4][1] = m.YV(jj + 1, kk + 1);
    movej = 0; movek = 0;
    for (int f = 0; f < 4; ++f) {
      double p[2], q[2], r[2];
      for (int n = 0; n < 2; ++n) {
        double dum = c[mo[f]][n];
        p[n] = c[ma[f]][n] - dum;
        q[n] = c[mb[f]][n] - dum;
        r[n] = xp[n] - dum;
      }
      double pqr = (p[0] * q[1] - p[1] * q[0]) * (p[0] * r[1] - p[1] * r[0]);
      if (pqr < 0.) {
        inside = false;
        if (f == 0) { --j; --movej; }
        if (f == 1) { --k; --movek; }
        if (f == 2) { ++j; ++movej; }
        if (f == 3) { ++k; ++movek; }
      }
    }
    bool alter = false;
    if ((j < 1 || j >= jmax) && movek != 0) { j -= movej; movej = 0; alter = true; }
    if ((k < 1 || k >= kmax) && movej != 0) { k -= movek; movek = 0; alter = true; }
    if (movejp + movej == 0 && movekp + movek == 0 && alter) { movej = 0; movek = 0; }
    if (movej == 0 && movek == 0 && !inside) outside = true;
    movejp = movej; movekp = movek;
    if (j < 1 || k < 1 || j >= jmax || k >= kmax) outside = true;
    if (nsearch >= 3) {
      if (3 * ((nsearch - 1) / 3) == nsearch - 1) { i1 = j; i2 = k; }
      else if (i1 == j && i2 == k) inside = true;
    }
    ++nsearch;
    if (nsearch > maxsearch) { outside = true; notoutside = true; }
  }
  if (inside && !outside) {
    js = j; ks = k;
  } else {
    js = -std::abs(j - (movej < 0 ? movej : 0));
    ks = -std::abs(k - (movek < 0 ? movek : 0));
    if (notoutside) ks = -ks;
  }
}

// -------------------------------------------------------- find_fractions
// 2x2 solve of matrixinv2x2 (ihc/src/ihc.f90:1463-1484): b(:,3) <- b(:,1:2)^-1 b(:,3)
void inv2(double b[2][3]) {
  double det = b[0][0] * b[1][1] - b[0][1] * b[1][0];
  double x = b[0][2] / det, y = b[1][2] / det;
  b[0][2] = x * b[1][1] - y * b[0][1];
  b[1][2] = -(x * b[1][0]) + y * b[0][0];
}

// fractions of xp in the 3x3 quadratic-Lagrange stencil of cell centres
// around (jd,kd) by Newton iteration (ihc/src/ihc.f90:1251-1430)
void find_fractions(const double xp[2], const MeshIn& m, int jd, int kd, double frac[2]) {
  const double onefourth = 1. / 4;
  double xv[9][2];
  int s = 0;
  for (int dk = -1; dk <= 1; ++dk)
    for (int dj = -1; dj <= 1; ++dj, ++s) {
      xv[s][0] = m.XC(jd + dj, kd + dk);
      xv[s][1] = m.YC(jd + dj, kd + dk);
    }
  double b[2][3];
  for (int n = 0; n < 2; ++n) {
    b[n][0] = 0.5 * (xv[5][n] - xv[3][n]);
    b[n][1] = 0.5 * (xv[7][n] - xv[1][n]);
    b[n][2] = xp[n] - xv[4][n];
  }
  inv2(b);
  for (int n = 0; n < 2; ++n) {
    frac[n] = b[n][2];
    if ((frac[n] - 1) * (frac[n] + 1) > 0.) frac[n] = .5;
  }
  int niter = 0;
  double resid = 1.e10;
  while (niter < 10 && resid > 1e-15) {
    double djm1 = 1. + frac[0], dkm1 = 1. + frac[1];
    double dj0 = frac[0], dk0 = frac[1];
    double djp1 = 1. - frac[0], dkp1 = 1. - frac[1];
    const double ddjm1 = 1., ddkm1 = 1., ddj0 = 1., ddk0 = 1., ddjp1 = -1., ddkp1 = -1.;
    double w[9], d1[9], d2[9];
    w[0] = (dj0 * djp1 * dk0 * dkp1) * onefourth;
    w[1] = -2 * (djm1 * djp1 * dk0 * dkp1) * onefourth;
    w[2] = -(djm1 * dj0 * dk0 * dkp1) * onefourth;
    w[3] = -2 * (dj0 * djp1 * dkm1 * dkp1) * onefourth;
    w[4] = 4 * (djm1 * djp1 * dkm1 * dkp1) * onefourth;
    w[5] = 2 * (djm1 * dj0 * dkm1 * dkp1) * onefourth;
    w[6] = -(dj0 * djp1 * dkm1 * dk0) * onefourth;
    w[7] = 2 * (djm1 * djp1 * dkm1 * dk0) * onefourth;
    w[8] = (djm1 * dj0 * dkm1 * dk0) * onefourth;
    d1[0] = ((ddj0 * djp1 * dk0 * dkp1) + (dj0 * ddjp1 * dk0 * dkp1)) * onefourth;
    d1[1] = -2 * ((ddjm1 * djp1 * dk0 * dkp1) + (djm1 * ddjp1 * dk0 * dkp1)) * onefourth;
    d1[2] = -((ddjm1 * dj0 * dk0 * dkp1) + (djm1 * ddj0 * dk0 * dkp1)) * onefourth;
    d1[3] = -2 * ((ddj0 * djp1 * dkm1 * dkp1) + (dj0 * ddjp1 * dkm1 * dkp1)) * onefourth;
    d1[4] = 4 * ((ddjm1 * djp1 * dkm1 * dkp1) + (djm1 * ddjp1 * dkm1 * dkp1)) * onefourth;
    d1[5] = 2 * ((ddjm1 * dj0 * dkm1 * dkp1) + (djm1 * ddj0 * dkm1 * dkp1)) * onefourth;
    d1[6] = -((ddj0 * djp1 * dkm1 * dk0) + (dj0 * ddjp1 * dkm1 * dk0)) * onefourth;
    d1[7] = 2 * ((ddjm1 * djp1 * dkm1 * dk0) + (djm1 * ddjp1 * dkm1 * dk0)) * onefourth;
    d1[8] = ((ddjm1 * dj0 * dkm1 * dk0) + (djm1 * ddj0 * dkm1 * dk0)) * onefourth;
    d2[0] = ((dj0 * djp1 * ddk0 * dkp1) + (dj0 * djp1 * dk0 * ddkp1)) * onefourth;
    d2[1] = -2 * ((djm1 * djp1 * ddk0 * dkp1) + (djm1 * djp1 * dk0 * ddkp1)) * onefourth;
    d2[2] = -((djm1 * dj0 * ddk0 * dkp1) + (djm1 * dj0 * dk0 * ddkp1)) * onefourth;
    d2[3] = -2 * ((dj0 * djp1 * ddkm1 * dkp1) + (dj0 * djp1 * dkm1 * ddkp1)) * onefourth;
    d2[4] = 4 * ((djm1 * djp1 * ddkm1 * dkp1) + (djm1 * djp1 * dkm1 * ddkp1)) * onefourth;
    d2[5] = 2 * ((djm1 * dj0 * ddkm1 * dkp1) + (djm1 * dj0 * dkm1 * ddkp1)) * onefourth;
    d2[6] = -((dj0 * djp1 * ddkm1 * dk0) + (dj0 * djp1 * dkm1 * ddk0)) * onefourth;
    d2[7] = 2 * ((djm1 * djp1 * ddkm1 * dk0) + (djm1 * djp1 * dkm1 * ddk0)) * onefourth;
    d2[8] = ((djm1 * dj0 * ddkm1 * dk0) + (djm1 * dj0 * dkm1 * ddk0)) * onefourth;
    for (int n = 0; n < 2; ++n) {
      double a1 = d1[0] * xv[0][n], a2 = d2[0] * xv[0][n], a3 = w[0] * xv[0][n];
      for (int i = 1; i < 9; ++i) {
        a1 = a1 + d1[i] * xv[i][n];
        a2 = a2 + d2[i] * xv[i][n];
        a3 = a3 + w[i] * xv[i][n];
      }
      b[n][0] = a1;
      b[n][1] = a2;
      b[n][2] = xp[n] - a3;
    }
    inv2(b);
    for (int n = 0; n < 2; ++n) {
      frac[n] = frac[n] + b[n][2];
      if ((frac[n] - 1.) * (frac[n] + 1.) > 0.) frac[n] = .5;
    }
    resid = 0.;
    for (int n = 0; n < 2; ++n) resid = resid + b[n][2] * b[n][2];
    ++niter;
  }
  if ((frac[0] + 1.0001) * (frac[0] - 1.0001) > 0. || (frac[1] + 1.0001) * (frac[1] - 1.0001) > 0.) {
    if (frac[0] < -1.0001) frac[0] = -0.99;
    if (frac[1] < -1.0001) frac[1] = -0.99;
    if (frac[0] > 1.0001) frac[0] = 0.99;
    if (frac[1] > 1.0001) frac[1] = 0.99;
  }
}

// -------------------------------------------------------- per-mesh state
struct Grid2 {
  int nj = 0, nk = 0;   // storage (0..nj-1) x (0..nk-1)
  std::vector<double> v;
  void init(int a, int b, double x) { nj = a; nk = b; v.assign((size_t)a * b, x); }
  double& operator()(int j, int k) { return v[(size_t)j + (size_t)k * nj]; }
};
struct Flags2 {
  int nj = 0, nk = 0;
  std::vector<uint8_t> v;
  void init(int a, int b) { nj = a; nk = b; v.assign((size_t)a * b, 0); }
  bool get(int j, int k) const { return v[(size_t)j + (size_t)k * nj] != 0; }
  void set(int j, int k) { v[(size_t)j + (size_t)k * nj] = 1; }
};

struct State {
  int jmaxc, kmaxc;
  Grid2 vold, volr;          // 0..jmax x 0..kmax
  std::vector<int> cell_bc;  // 0..jmax x 0..kmax
  Flags2 wall;               // 1..jmax x 1..kmax (stored 0..jmax)
  Flags2 forcerecv, noblank, ichk;
  int& bc(int j, int k) { return cell_bc[(size_t)j + (size_t)k * vold.nj]; }
};

struct Hole {
  int j, k, n, js, ks;
  double f[2];
};

// apply the rows of one ihcbc file (read_bc, ihc/src/ihc.f90:147-208)
void read_bc(const MeshIn& m, State& s) {
  const int jmax = m.jmax, kmax = m.kmax;
  for (int ib = 0; ib < m.nbc; ++ib) {
    int js = m.jbcs[ib], je = m.jbce[ib], ks = m.kbcs[ib], ke = m.kbce[ib];
    if (m.ibtyp[ib] == 5) {
      if (js < 0) js = jmax + js + 1;
      if (ks < 0) ks = kmax + ks + 1;
      if (je < 0) je = jmax + je + 1;
      if (ke < 0) ke = kmax + ke + 1;
      for (int j = js; j <= je; ++j)
        for (int k = ks; k <= ke; ++k) s.wall.set(j, k);
    } else {
      if (js < 0) js = s.jmaxc + js + 1;
      if (ks < 0) ks = s.kmaxc + ks + 1;
      if (je < 0) je = s.jmaxc + je + 1;
      if (ke < 0) ke = s.kmaxc + ke + 1;
      for (int j = js; j <= je; ++j)
        for (int k = ks; k <= ke; ++k) s.bc(j, k) = m.ibtyp[ib];
    }
  }
}

struct Ctx {
  int ngrids;
  const MeshIn* m;
  std::vector<Octree> tree;
  std::vector<State> st;
  std::vector<std::vector<int>> iblank;   // 0..jmax x 0..kmax per mesh
  std::vector<std::vector<uint8_t>> overlap;
  std::vector<std::vector<Hole>> hole;
  int& ib(int n, int j, int k) { return iblank[n][(size_t)j + (size_t)k * (m[n].jmax + 1)]; }
};

// the octree-guided restarts of a failed walk (ihc/src/ihc.f90:531-585,
// 687-758): up to ntmax rounds over the start and end vertices of the point's
// box at its deepest non-empty level (and the level above it); nbodycross
// counts walks that ended on a wall point
void restart_search(Ctx& c, int n, const double xp[2], const double xbar[2], int& js, int& ks, int ntmax,
                    int& nlvl, int* nbodycross) {
  const MeshIn& md = c.m[n];
  const Octree& t = c.tree[n];
  int lvlcheck = t.lvlpoint[box_index(xbar, kLvlMax)];
  nlvl = 2;
  if (lvlcheck == kLvlMin) nlvl = 1;
  int jst[2], kst[2], jet[2], ket[2];
  for (int l = 1; l <= nlvl; ++l) {
    int64_t nout = box_index(xbar, lvlcheck - l + 1);
    jst[l - 1] = t.istart[0][nout]; kst[l - 1] = t.istart[1][nout];
    jet[l - 1] = t.iend[0][nout];   ket[l - 1] = t.iend[1][nout];
  }
  auto crossed = [&](int a, int b) {
    if (!nbodycross) return;
    int ja = std::abs(a), kb = std::abs(b);
    if (ja >= 1 && ja <= md.jmax && kb >= 1 && kb <= md.kmax && c.st[n].wall.get(ja, kb)) ++*nbodycross;
  };
  int nstop = 0, ntime = 0;
  while (nstop < 2 * nlvl && ntime < ntmax) {
    for (int l = 0; l < nlvl; ++l) {
      if (js <= 0 && kst[l] > 0) {
        js = std::abs(jst[l]); ks = std::abs(kst[l]);
        finddonor(xp, md, js, ks, 100);
        jst[l] = js; kst[l] = ks;
        if (js <= 0 && ks <= 0) { crossed(js, ks); ++nstop; }
      } else {
        ++nstop;
      }
      if (js <= 0 && ket[l] > 0) {
        js = std::abs(jet[l]); ks = std::abs(ket[l]);
        finddonor(xp, md, js, ks, 100);
        jet[l] = js; ket[l] = ks;
        if (js <= 0 && ks <= 0) { crossed(js, ks); ++nstop; }
      } else {
        ++nstop;
      }
    }
    ++ntime;
  }
}

// boundconnect (ihc/src/ihc.f90:451-614): does some boundary cell of mesh
// ng lie inside mesh n?
void boundconnect(Ctx& c, int ng) {
  const MeshIn& mg = c.m[ng];
  const int jmaxc = mg.jmax - 1, kmaxc = mg.kmax - 1;
  for (int n = 0; n < c.ngrids; ++n) {
    if (c.overlap[n][ng]) c.overlap[ng][n] = 1;
    if (n == ng || c.overlap[ng][n]) continue;
    const Octree& t = c.tree[n];
    int jl = 2, kl = 2, js = -1, ks = -1;
    int j = 1, k = 1, jstep = 1, kstep = 1, nface = 1;
    while (true) {
      if (nface == 1) j = 1;
      if (nface == 2) j = jmaxc;
      if (nface == 3) k = 1;
      if (nface == 4) k = kmaxc;
      double xp[2] = {mg.XC(j, k), mg.YC(j, k)};
      double xbar[2] = {(xp[0] - t.xs[0]) / t.scale[0], (xp[1] - t.xs[1]) / t.scale[1]};
      bool skip = true;
      if (xbar[0] >= 0 && xbar[1] >= 0 && xbar[0] < 1 && xbar[1] < 1) {
        if (!t.empty[box_index(xbar, kLvlMin)]) {
          skip = false;
          js = jl; ks = kl;
          finddonor(xp, c.m[n], js, ks, 100);
          if (js <= 0) {
            int nlvl;
            restart_search(c, n, xp, xbar, js, ks, 1, nlvl, nullptr);
          }
          if (js > 0) {
            jl = js; kl = ks;
            c.overlap[ng][n] = 1;
            c.overlap[n][ng] = 1;
            break;
          }
        }
      }
      (void)skip;
      if (nface == 1 || nface == 3) {
        k += kstep;
        if (k > kmaxc || k < 1) { ++nface; kstep = -kstep; k += kstep; continue; }
      }
      if (nface == 2 || nface == 4) {
        j += jstep;
        if (j > jmaxc || j < 1) { ++nface; jstep = -jstep; j += jstep; continue; }
      }
      if (nface == 5) break;
    }
  }
}

// recvconnect (forced receivers, ihc/src/ihc.f90:617-865) and holeconnect
// (every other cell, 868-1115) of mesh ng
void connect_points(Ctx& c, int ng, bool forced) {
  const MeshIn& mg = c.m[ng];
  State& sg = c.st[ng];
  const int jmaxc = sg.jmaxc, kmaxc = sg.kmaxc;
  std::vector<double> volrecv((size_t)(jmaxc + 1) * (kmaxc + 1));
  auto VR = [&](int j, int k) -> double& { return volrecv[(size_t)j + (size_t)k * (jmaxc + 1)]; };
  for (int j = 1; j <= jmaxc; ++j)
    for (int k = 1; k <= kmaxc; ++k) VR(j, k) = kOverlap * sg.volr(j, k);
  std::vector<int> ptr((size_t)(jmaxc + 1) * (kmaxc + 1), 0);
  auto PTR = [&](int j, int k) -> int& { return ptr[(size_t)j + (size_t)k * (jmaxc + 1)]; };
  std::vector<Hole>& holes = c.hole[ng];
  const size_t first_new = holes.size();
  int nlvl = 0;
  for (int n = 0; n < c.ngrids; ++n) {
    if (n == ng || !c.overlap[ng][n]) continue;
    const MeshIn& md = c.m[n];
    State& sd = c.st[n];
    const Octree& t = c.tree[n];
    int jl = 2, kl = 2, js = -1, ks = -1;
    int j = 1, k = 1, jstep = 1, kstep = 1;
    const double si0 = 1. / t.scale[0], si1 = 1. / t.scale[1];
    while (true) {
      int nbodycross = 0;
      bool consider = forced ? sg.forcerecv.get(j, k) : !(VR(j, k) < 0. || sg.ichk.get(j, k));
      if (consider) {
        double xp[2] = {mg.XC(j, k), mg.YC(j, k)};
        double xbar[2] = {(xp[0] - t.xs[0]) * si0, (xp[1] - t.xs[1]) * si1};
        bool inbox = xbar[0] >= 0 && xbar[1] >= 0 && xbar[0] < 1 && xbar[1] < 1;
        if (forced) inbox = VR(j, k) >= 0. && inbox;
        if (inbox && !t.empty[box_index(xbar, kLvlMin)]) {
          js = jl; ks = kl;
          finddonor(xp, md, js, ks, 100);
          if (js <= 0) {
            if (ks <= 0) {
              int ja = std::abs(js), kb = std::abs(ks);
              if (ja >= 1 && ja <= md.jmax && kb >= 1 && kb <= md.kmax && sd.wall.get(ja, kb)) ++nbodycross;
            }
            restart_search(c, n, xp, xbar, js, ks, 5, nlvl, &nbodycross);
          }
          if (js > 0) {
            double voldonor = sd.vold(js, ks);
            bool fr = false;
            for (int dk = -1; dk <= 1; ++dk)
              for (int dj = -1; dj <= 1; ++dj) fr = fr || sd.forcerecv.get(js + dj, ks + dk);
            if (fr) voldonor = 1.e30;
            if (voldonor < VR(j, k)) {
              VR(j, k) = voldonor;
              size_t nc;
              if (PTR(j, k) == 0) {
                holes.push_back(Hole{});
                nc = holes.size() - 1;
                PTR(j, k) = (int)(nc + 1);
              } else {
                nc = (size_t)PTR(j, k) - 1;
              }
              Hole& h = holes[nc];
              h.j = j; h.k = k; h.n = n; h.js = js; h.ks = ks;
              for (int dk = -1; dk <= 1; ++dk)
                for (int dj = -1; dj <= 1; ++dj) sg.vold(j + dj, k + dk) = 1.e30;
              find_fractions(xp, md, js, ks, h.f);
            }
            jl = js; kl = ks;
          }
          if (js <= 0 && nbodycross > 2 * nlvl) {
            if (!sg.noblank.get(j, k)) {
              c.ib(ng, j, k) = 0;
              VR(j, k) = -1.e30;
            }
          }
          if (forced) sg.ichk.set(j, k);
        }
      }
      k += kstep;
      if (k > kmaxc || k < 1) { kstep = -kstep; k += kstep; j += jstep; }
      if (j > jmaxc) break;
    }
  }
  // every listed donor's neighbourhood stops receiving (volr = 0)
  for (size_t nc = 0; nc < holes.size(); ++nc) {
    (void)first_new;
    const Hole& h = holes[nc];
    State& sd = c.st[h.n];
    for (int dk = -1; dk <= 1; ++dk)
      for (int dj = -1; dj <= 1; ++dj) sd.volr(h.js + dj, h.ks + dk) = 0.;
  }
}

}  // namespace

int connect(const MeshIn* meshes, int ngrids, int nhalo, Result* out) {
  Ctx c;
  c.ngrids = ngrids;
  c.m = meshes;
  c.tree.resize(ngrids);
  c.st.resize(ngrids);
  c.iblank.resize(ngrids);
  c.hole.resize(ngrids);
  c.overlap.assign(ngrids, std::vector<uint8_t>(ngrids, 0));
  for (int n = 0; n < ngrids; ++n) {
    const MeshIn& m = meshes[n];
    if (m.jmax < 3 || m.kmax < 3) return -1;
    State& s = c.st[n];
    s.jmaxc = m.jmax - 1; s.kmaxc = m.kmax - 1;
    s.vold.init(m.jmax + 1, m.kmax + 1, 0.);
    s.volr.init(m.jmax + 1, m.kmax + 1, 0.);
    s.cell_bc.assign((size_t)(m.jmax + 1) * (m.kmax + 1), 0);
    s.wall.init(m.jmax + 1, m.kmax + 1);
    s.forcerecv.init(m.jmax + 1, m.kmax + 1);
    s.noblank.init(m.jmax + 1, m.kmax + 1);
    s.ichk.init(m.jmax + 1, m.kmax + 1);
    read_bc(m, s);
    for (int k = 0; k <= m.kmax; ++k)
      for (int j = 0; j <= m.jmax; ++j) { s.vold(j, k) = m.VOL(j, k); s.volr(j, k) = m.VOL(j, k); }
    for (int j = 1; j <= s.jmaxc; ++j)
      for (int k = 1; k <= s.kmaxc; ++k) {
        int b = s.bc(j, k);
        if (b == 1 || b == 6) { s.volr(j, k) = 0.; s.vold(j, k) = 0.; }
        if (b == 2) { s.volr(j, k) = -1.e30; s.vold(j, k) = 1.e30; }
        if (b == 3 || b == 7) { s.volr(j, k) = 1.e30; s.vold(j, k) = 1.e30; s.forcerecv.set(j, k); }
        if (b == 4 || b == 6 || b == 7) s.noblank.set(j, k);
      }
    make_octree(m, c.tree[n]);
    c.iblank[n].assign((size_t)(m.jmax + 1) * (m.kmax + 1), 1);
    for (int j = 1; j <= s.jmaxc; ++j)
      for (int k = 1; k <= s.kmaxc; ++k)
        if (s.bc(j, k) == 2) c.ib(n, j, k) = 0;
  }
  for (int n = 0; n < ngrids; ++n) boundconnect(c, n);
  for (int n = 0; n < ngrids; ++n) connect_points(c, n, true);
  for (int n = 0; n < ngrids; ++n) connect_points(c, n, false);
  // lists (ihc/src/ihc.f90:390-446)
  out->mesh.assign(ngrids, MeshOut{});
  std::vector<std::vector<int>> ibctmp(ngrids), idontmp(ngrids);
  for (int n = 0; n < ngrids; ++n) {
    for (const Hole& h : c.hole[n]) {
      if (c.ib(n, h.j, h.k) == 0) continue;
      MeshOut& r = out->mesh[n];
      MeshOut& d = out->mesh[h.n];
      r.imesh.push_back(h.j + nhalo);
      r.imesh.push_back(h.k + nhalo);
      d.idonor.push_back(h.js + nhalo);
      d.idonor.push_back(h.ks + nhalo);
      d.frac.push_back(h.f[0]);
      d.frac.push_back(h.f[1]);
      ibctmp[n].push_back((int)(d.idonor.size() / 2));
      idontmp[n].push_back(h.n);
      c.ib(n, h.j, h.k) = -1;
    }
  }
  std::vector<int> ndon(ngrids);
  for (int n = 0; n < ngrids; ++n) ndon[n] = (int)(out->mesh[n].idonor.size() / 2);
  int sp = 1;
  for (int n = 0; n < ngrids; ++n) {
    out->mesh[n].iisptr = sp;
    out->mesh[n].iieptr = sp + ndon[n] - 1;
    sp += ndon[n];
  }
  for (int n = 0; n < ngrids; ++n) {
    MeshOut& r = out->mesh[n];
    for (size_t i = 0; i < ibctmp[n].size(); ++i) {
      int count = 0;
      for (int nn = 0; nn < idontmp[n][i]; ++nn) count += ndon[nn];
      r.ibc.push_back(count + ibctmp[n][i]);
    }
    // iblank of the solver's cells: IHC cell (j,k) is solver cell (j+1,k+1)
    r.iblank = c.iblank[n];
  }
  return 0;
}

}  // namespace ihc
}  // namespace ot2d
