// gfx950 kernels of the viscous residual and the Spalart-Allmaras model.
//
//   visrhs (src/visrhs.f:2-287)        k_visrhs      one workgroup per 32x8 tile of cells:
//                                                    per-cell primitives, coefficients and
//                                                    cross terms once into LDS, then each
//                                                    cell's thin-layer fluxes from there
//   vmu_sa (src/samodel2d.f:2-197)     k_sa_prim     laminar viscosity, u, v (all cells)
//                                      k_turbc       SA boundary conditions (1 block / segment,
//                                                    one launch per phase of the segments)
//                                      k_turbc_eddy  the final SA boundary conditions and the
//                                                    eddy viscosity of the halo ring
//                                      k_sa_rhs      convection, diffusion, source and the
//                                                    two scalar tridiagonal systems, in
//                                                    gather form (one thread per cell)
//                                      k_sa_solve_j  tridiagonal solve along j from both ends
//                                                    (two lanes per k row)
//                                      k_sa_solve_kt tridiagonal solves along k, 16 or 8 adjacent
//                                                    j columns per workgroup (k_sa_solve_k: one
//                                                    column, for lines over 256 cells),
//                                                    nu-tilde update and eddy viscosity
//
// Operand order follows the reference expressions (built with
// -ffp-contract=off), so results agree with the reference to the rounding of
// the device's sqrt/div/pow; the line solves of lines over 32 cells run from
// both ends (thomas2_lds), which changes their rounding only.
#include "ot2d_dev.h"

namespace ot2d {

__device__ __forceinline__ double sqd(double a) { return a * a; }
__device__ __forceinline__ int imx(int a, int b) { return a > b ? a : b; }

// ====================================================================== visrhs
struct VisPrim {
  double u, v, e, vnu, gkap;
};

// per-cell quantities of the visrhs line loops (src/visrhs.f:47-57)
__device__ __forceinline__ VisPrim vis_prim_of(double q1, double q2, double q3, double q4, double turm, double ggm1,
                                               double c2b, double c2bp, double prtr) {
  VisPrim p;
  double ra = 1. / q1;
  p.u = q2 * ra;
  p.v = q3 * ra;
  p.e = q4 * ra - .5 * (sqd(p.u) + sqd(p.v));
  double tt = ggm1 * p.e;
  double vmue = c2bp * tt * sqrt(tt) / (c2b + tt);
  p.vnu = vmue + turm;
  p.gkap = vmue + prtr * turm;
  return p;
}
__device__ __forceinline__ VisPrim vis_prim(const DevMesh& M, long long i, double c2b, double c2bp, double prtr) {
  return vis_prim_of(M.q[0][i], M.q[1][i], M.q[2][i], M.q[3][i], M.turmu[i], M.c.ggm1, c2b, c2bp, prtr);
}

struct VisCoef {
  double c1, c2, c5, c4;          // thin-layer coefficients (b1v.. or a1v..)
  double d1, d2, d4, d5, d7;      // cross-derivative coefficients
};

// metrics of a cell and the reciprocal of its Jacobian q(.,.,nq)
struct VisMet {
  double xx, xy, yx, yy, r;
};
__device__ __forceinline__ VisMet vis_met(const DevMesh& M, long long i) {
  VisMet m;
  m.xx = M.xx[i];
  m.xy = M.xy[i];
  m.yx = M.yx[i];
  m.yy = M.yy[i];
  m.r = 1. / M.q[M.nq - 1][i];
  return m;
}

// coefficients of one line direction at a cell: the thin-layer terms from
// the cell's metric pair (ma, mb) = (yx, yy) along eta, (xx, xy) along xi,
// and the cross terms from its two neighbours m, p across the line
// (src/visrhs.f:59-92 eta, 184-217 xi; both use the same expressions)
__device__ __forceinline__ VisCoef vis_coef(double ma, double mb, double rj, const VisMet& m, const VisMet& p,
                                            const VisPrim& c, double dre, double gkpr) {
  VisCoef o;
  double b4 = (sqd(ma) + sqd(mb)) * rj;
  o.c1 = (b4 + sqd(ma) / 3. * rj) * c.vnu * dre;
  o.c2 = (b4 + sqd(mb) / 3. * rj) * c.vnu * dre;
  o.c5 = (ma * mb / 3. * rj) * c.vnu * dre;
  o.c4 = b4 * gkpr * c.gkap * dre;
  double t;
  t = (4. / 3. * m.xx * m.yx + m.xy * m.yy) * m.r;
  t = t + (4. / 3. * p.xx * p.yx + p.xy * p.yy) * p.r;
  o.d1 = c.vnu * dre * t;
  t = (4. / 3. * m.xy * m.yy + m.xx * m.yx) * m.r;
  t = t + (4. / 3. * p.xy * p.yy + p.xx * p.yx) * p.r;
  o.d2 = c.vnu * dre * t;
  t = (m.xy * m.yy + m.xx * m.yx) * m.r;
  t = t + (p.xy * p.yy + p.xx * p.yx) * p.r;
  o.d4 = t * gkpr * c.gkap * dre;
  t = (-2. / 3. * m.xy * m.yx + m.xx * m.yy) * m.r;
  t = t + (-2. / 3. * p.xy * p.yx + p.xx * p.yy) * p.r;
  o.d5 = c.vnu * dre * t;
  t = (-2. / 3. * m.xx * m.yy + m.xy * m.yx) * m.r;
  t = t + (-2. / 3. * p.xx * p.yy + p.xy * p.yx) * p.r;
  o.d7 = c.vnu * dre * t;
  return o;
}

// eta direction (j = const line) coefficients of cell (j,k) from global memory
__device__ __forceinline__ VisCoef eta_coef(const DevMesh& M, int j, int k, const VisPrim& p, double dre,
                                            double gkpr) {
  long long i = cell(M, j, k);
  return vis_coef(M.yx[i], M.yy[i], 1. / M.q[M.nq - 1][i], vis_met(M, cell(M, j - 1, k)),
                  vis_met(M, cell(M, j + 1, k)), p, dre, gkpr);
}

// xi direction (k = const line) coefficients of cell (j,k) from global memory
__device__ __forceinline__ VisCoef xi_coef(const DevMesh& M, int j, int k, const VisPrim& p, double dre,
                                           double gkpr) {
  long long i = cell(M, j, k);
  return vis_coef(M.xx[i], M.xy[i], 1. / M.q[M.nq - 1][i], vis_met(M, cell(M, j, k - 1)),
                  vis_met(M, cell(M, j, k + 1)), p, dre, gkpr);
}

struct Tri { double f2, f3, f4; };

// thin-layer flux across the face between cells a (low) and b (high) of a line
__device__ __forceinline__ Tri vis_face(const VisPrim& pa, const VisCoef& ca, const VisPrim& pb, const VisCoef& cb) {
  double b1 = cb.c1 + ca.c1, b2 = cb.c2 + ca.c2, b5 = cb.c5 + ca.c5;
  double t4 = pb.u * cb.c1 + pa.u * ca.c1;
  double t5 = pb.v * cb.c2 + pa.v * ca.c2;
  double t6 = pb.u * cb.c5 + pa.u * ca.c5;
  double t7 = pb.v * cb.c5 + pa.v * ca.c5;
  double t8 = cb.c4 + ca.c4;
  double du = pb.u - pa.u, dv = pb.v - pa.v, dei = pb.e - pa.e;
  Tri f;
  f.f2 = b1 * du + b5 * dv;
  f.f3 = b5 * du + b2 * dv;
  f.f4 = (t4 + t7) * du + (t5 + t6) * dv + t8 * dei;
  return f;
}

// velocities and internal energy by division (q2/q1 ...), as the cross
// derivative terms use them
struct DivPrim { double u, v, e; };
__device__ __forceinline__ DivPrim div_prim_of(double q1, double q2, double q3, double q4) {
  DivPrim d;
  d.u = q2 / q1;
  d.v = q3 / q1;
  d.e = q4 / q1 - .5 * (sqd(d.u) + sqd(d.v));
  return d;
}
__device__ __forceinline__ DivPrim div_prim(const DevMesh& M, long long i) {
  return div_prim_of(M.q[0][i], M.q[1][i], M.q[2][i], M.q[3][i]);
}

// cross-derivative term of the eta line at cell c from its j neighbours
// (src/visrhs.f:106-146)
__device__ __forceinline__ Tri eta_cross(const VisCoef& c, const DivPrim& p, const DivPrim& m) {
  double duj = 0.5 * (p.u - m.u), dvj = 0.5 * (p.v - m.v), dej = 0.5 * (p.e - m.e);
  Tri f;
  f.f2 = c.d1 * duj + c.d5 * dvj;
  f.f3 = c.d7 * duj + c.d2 * dvj;
  double t = 0.5 * c.d1 * 0.5 * (sqd(p.u) - sqd(m.u));
  t = t + 0.5 * c.d2 * 0.5 * (sqd(p.v) - sqd(m.v));
  t = t + c.d5 * 0.5 * (p.u + m.u) * dvj;
  t = t + c.d7 * 0.5 * (p.v + m.v) * duj;
  f.f4 = t + c.d4 * dej;
  return f;
}

// cross-derivative term of the xi line (src/visrhs.f:227-268)
__device__ __forceinline__ Tri xi_cross(const VisCoef& c, const DivPrim& p, const DivPrim& m) {
  double duk = 0.5 * (p.u - m.u), dvk = 0.5 * (p.v - m.v), dek = 0.5 * (p.e - m.e);
  Tri f;
  f.f2 = c.d1 * duk + c.d7 * dvk;
  f.f3 = c.d5 * duk + c.d2 * dvk;
  double t = 0.5 * c.d1 * 0.5 * (sqd(p.u) - sqd(m.u));
  t = t + 0.5 * c.d2 * 0.5 * (sqd(p.v) - sqd(m.v));
  t = t + c.d5 * 0.5 * (p.v + m.v) * duk;
  t = t + c.d7 * 0.5 * (p.u + m.u) * dvk;
  f.f4 = t + c.d4 * dek;
  return f;
}

// A workgroup owns VIS_TJ x VIS_TK interior cells.  Every per-cell quantity
// of the reference's line loops -- primitives, quotient primitives, metrics,
// the coefficients and cross terms of both directions -- is formed once for
// the cells of the tile and of its one-cell ring and kept in LDS; each cell
// then assembles its two face fluxes per direction from there.  The
// arithmetic of every quantity is the reference's, so the result does not
// depend on the tiling.
constexpr int VIS_TJ = 32, VIS_TK = 8, VIS_THREADS = VIS_TJ * VIS_TK;
constexpr int VIS_RJ = VIS_TJ + 2, VIS_RK = VIS_TK + 2;
constexpr int VIS_NR = VIS_RJ * VIS_RK;   // tile and ring, ring row 0 = k0-1, column 0 = j0-1
constexpr int VIS_NE = VIS_TJ * VIS_RK;   // eta cells: tile columns, rows k0-1 .. k0+TK
constexpr int VIS_NX = VIS_RJ * VIS_TK;   // xi cells: columns j0-1 .. j0+TJ, tile rows

struct VisLds {
  double u[VIS_NR], v[VIS_NR], e[VIS_NR], vnu[VIS_NR], gkap[VIS_NR];
  double du[VIS_NR], dv[VIS_NR], de[VIS_NR];
  double xx[VIS_NR], xy[VIS_NR], yx[VIS_NR], yy[VIS_NR], r[VIS_NR];
  double ec[4][VIS_NE], eg[3][VIS_NE];     // eta: c1 c2 c5 c4 and the cross term
  double xc[4][VIS_NX], xg[3][VIS_NX];     // xi
};
constexpr size_t VIS_LDS = sizeof(VisLds);

__device__ __forceinline__ VisPrim lds_prim(const VisLds& L, int c) {
  VisPrim p;
  p.u = L.u[c]; p.v = L.v[c]; p.e = L.e[c]; p.vnu = L.vnu[c]; p.gkap = L.gkap[c];
  return p;
}
__device__ __forceinline__ DivPrim lds_div(const VisLds& L, int c) {
  DivPrim d;
  d.u = L.du[c]; d.v = L.dv[c]; d.e = L.de[c];
  return d;
}
__device__ __forceinline__ VisMet lds_met(const VisLds& L, int c) {
  VisMet m;
  m.xx = L.xx[c]; m.xy = L.xy[c]; m.yx = L.yx[c]; m.yy = L.yy[c]; m.r = L.r[c];
  return m;
}
template <int N>
__device__ __forceinline__ VisCoef lds_coef(const double (&cf)[4][N], int c) {
  VisCoef o{};
  o.c1 = cf[0][c]; o.c2 = cf[1][c]; o.c5 = cf[2][c]; o.c4 = cf[3][c];
  return o;
}
template <int N>
__device__ __forceinline__ Tri lds_tri(const double (&g)[3][N], int c) {
  Tri f;
  f.f2 = g[0][c]; f.f3 = g[1][c]; f.f4 = g[2][c];
  return f;
}

// With norms set, the residual is complete once this kernel has added the
// viscous terms, so each tile also takes its residual-norm partials here
// (the tiles and thread-to-cell map are the residual's: 32x8 tiles, k-tile
// major), instead of a separate pass over s.
__global__ __launch_bounds__(VIS_THREADS) void k_visrhs(const DevMesh* __restrict__ Ms, int norms) {
  static_assert(VIS_TJ == RES_TJ && VIS_TK == RES_TK, "visrhs tiles are the residual's");
  __shared__ double red[4][RES_THREADS];
  const DevMesh& M = Ms[blockIdx.y];
  const int nj = M.je - M.js + 1, nk = M.ke - M.ks + 1;
  const int tiles_j = (nj + VIS_TJ - 1) / VIS_TJ;
  if ((int)blockIdx.x >= tiles_j * ((nk + VIS_TK - 1) / VIS_TK) || *M.status) return;
  extern __shared__ __attribute__((aligned(16))) double vis_raw[];
  VisLds& L = *reinterpret_cast<VisLds*>(vis_raw);
  const DevConst& C = M.c;
  const double gkpr = C.gamma / C.pr, prtr = C.pr / 0.9, dre = .5 / C.rey;
  const double c2b = 198.6 / C.tinf, c2bp = c2b + 1.;
  const int j0 = M.js + ((int)blockIdx.x % tiles_j) * VIS_TJ, k0 = M.ks + ((int)blockIdx.x / tiles_j) * VIS_TK;
  const int t = threadIdx.x;
  const double* qj = M.q[M.nq - 1];

  // tile and ring: primitives (src/visrhs.f:47-57), quotient primitives
  // (src/visrhs.f:108-119), metrics and 1/q(nq)
  for (int n = t; n < VIS_NR; n += VIS_THREADS) {
    const int j = j0 - 1 + n % VIS_RJ, k = k0 - 1 + n / VIS_RJ;
    if (j > M.je + 1 || k > M.ke + 1) continue;
    const long long i = cell(M, j, k);
    const double q1 = M.q[0][i], q2 = M.q[1][i], q3 = M.q[2][i], q4 = M.q[3][i], turm = M.turmu[i];
    const double xx = M.xx[i], xy = M.xy[i], yx = M.yx[i], yy = M.yy[i], qn = qj[i];
    const VisPrim p = vis_prim_of(q1, q2, q3, q4, turm, C.ggm1, c2b, c2bp, prtr);
    const DivPrim d = div_prim_of(q1, q2, q3, q4);
    L.u[n] = p.u; L.v[n] = p.v; L.e[n] = p.e; L.vnu[n] = p.vnu; L.gkap[n] = p.gkap;
    L.du[n] = d.u; L.dv[n] = d.v; L.de[n] = d.e;
    L.xx[n] = xx; L.xy[n] = xy; L.yx[n] = yx; L.yy[n] = yy; L.r[n] = 1. / qn;
  }
  __syncthreads();

  // coefficients and cross terms: eta along the tile columns for rows
  // k0-1 .. k0+TK, xi along the tile rows for columns j0-1 .. j0+TJ
  for (int n = t; n < VIS_NE + VIS_NX; n += VIS_THREADS) {
    if (n < VIS_NE) {
      const int a = n % VIS_TJ, b = n / VIS_TJ;
      if (j0 + a > M.je || k0 - 1 + b > M.ke + 1) continue;
      const int c = b * VIS_RJ + a + 1;
      const VisCoef o = vis_coef(L.yx[c], L.yy[c], L.r[c], lds_met(L, c - 1), lds_met(L, c + 1), lds_prim(L, c), dre,
                                 gkpr);
      const Tri g = eta_cross(o, lds_div(L, c + 1), lds_div(L, c - 1));
      L.ec[0][n] = o.c1; L.ec[1][n] = o.c2; L.ec[2][n] = o.c5; L.ec[3][n] = o.c4;
      L.eg[0][n] = g.f2; L.eg[1][n] = g.f3; L.eg[2][n] = g.f4;
    } else {
      const int m = n - VIS_NE, a = m % VIS_RJ, b = m / VIS_RJ;
      if (j0 - 1 + a > M.je + 1 || k0 + b > M.ke) continue;
      const int c = (b + 1) * VIS_RJ + a;
      const VisCoef o = vis_coef(L.xx[c], L.xy[c], L.r[c], lds_met(L, c - VIS_RJ), lds_met(L, c + VIS_RJ),
                                 lds_prim(L, c), dre, gkpr);
      const Tri g = xi_cross(o, lds_div(L, c + VIS_RJ), lds_div(L, c - VIS_RJ));
      L.xc[0][m] = o.c1; L.xc[1][m] = o.c2; L.xc[2][m] = o.c5; L.xc[3][m] = o.c4;
      L.xg[0][m] = g.f2; L.xg[1][m] = g.f3; L.xg[2][m] = g.f4;
    }
  }
  __syncthreads();

  const int a = t % VIS_TJ, b = t / VIS_TJ;
  const int j = j0 + a, k = k0 + b;
  const bool valid = j <= M.je && k <= M.ke;
  // every thread reaches the one norm_tile call below (its barriers), so
  // cells outside the grid skip the work instead of returning
  double sv[4] = {0.0, 0.0, 0.0, 0.0};
  if (valid) {
    const long long i = cell(M, j, k);
    double s2 = M.s[1][i], s3 = M.s[2][i], s4 = M.s[3][i];
    const int c = (b + 1) * VIS_RJ + a + 1;
    const VisPrim p1 = lds_prim(L, c);
    // ---- eta part (j = const line): faces k-1|k, k|k+1, cross terms
    {
      const int e = (b + 1) * VIS_TJ + a;
      const VisCoef c0 = lds_coef(L.ec, e - VIS_TJ), c1 = lds_coef(L.ec, e), c2 = lds_coef(L.ec, e + VIS_TJ);
      const Tri fm = vis_face(lds_prim(L, c - VIS_RJ), c0, p1, c1);   // face k-1
      const Tri fp = vis_face(p1, c1, lds_prim(L, c + VIS_RJ), c2);   // face k
      const Tri gm = lds_tri(L.eg, e - VIS_TJ), gc = lds_tri(L.eg, e);
      if (k < M.ke) {
        const Tri gp = lds_tri(L.eg, e + VIS_TJ);
        s2 = s2 + (fp.f2 - fm.f2) + 0.5 * (gp.f2 - gm.f2);
        s3 = s3 + (fp.f3 - fm.f3) + 0.5 * (gp.f3 - gm.f3);
        s4 = s4 + (fp.f4 - fm.f4) + 0.5 * (gp.f4 - gm.f4);
      } else {  // one-sided closure at kend (src/visrhs.f:141-150)
        Tri gq;
        if (b >= 1) {
          gq = lds_tri(L.eg, e - 2 * VIS_TJ);
        } else {  // row k-2 lies below the tile's ring
          const VisPrim pq = vis_prim(M, cell(M, j, k - 2), c2b, c2bp, prtr);
          const VisCoef cq = eta_coef(M, j, k - 2, pq, dre, gkpr);
          gq = eta_cross(cq, div_prim(M, cell(M, j + 1, k - 2)), div_prim(M, cell(M, j - 1, k - 2)));
        }
        s2 = s2 + (fp.f2 - fm.f2) + 0.5 * (gq.f2 - 4. * gm.f2 + 3 * gc.f2);
        s3 = s3 + (fp.f3 - fm.f3) + 0.5 * (gq.f3 - 4. * gm.f3 + 3 * gc.f3);
        s4 = s4 + (fp.f4 - fm.f4) + 0.5 * (gq.f4 - 4. * gm.f4 + 3. * gc.f4);
      }
    }
    // ---- xi part (k = const line)
    {
      const int x = b * VIS_RJ + a + 1;
      const VisCoef c0 = lds_coef(L.xc, x - 1), c1 = lds_coef(L.xc, x), c2 = lds_coef(L.xc, x + 1);
      const Tri fm = vis_face(lds_prim(L, c - 1), c0, p1, c1);
      const Tri fp = vis_face(p1, c1, lds_prim(L, c + 1), c2);
      const Tri gm = lds_tri(L.xg, x - 1), gc = lds_tri(L.xg, x);
      if (j < M.je) {
        const Tri gp = lds_tri(L.xg, x + 1);
        s2 = s2 + (fp.f2 - fm.f2) + 0.5 * (gp.f2 - gm.f2);
        s3 = s3 + (fp.f3 - fm.f3) + 0.5 * (gp.f3 - gm.f3);
        s4 = s4 + (fp.f4 - fm.f4) + 0.5 * (gp.f4 - gm.f4);
      } else {  // one-sided closure at jend (src/visrhs.f:262-271)
        Tri gq;
        if (a >= 1) {
          gq = lds_tri(L.xg, x - 2);
        } else {  // column j-2 lies left of the tile's ring
          const VisPrim pq = vis_prim(M, cell(M, j - 2, k), c2b, c2bp, prtr);
          const VisCoef cq = xi_coef(M, j - 2, k, pq, dre, gkpr);
          gq = xi_cross(cq, div_prim(M, cell(M, j - 2, k + 1)), div_prim(M, cell(M, j - 2, k - 1)));
        }
        s2 = s2 + (fp.f2 - fm.f2) + 0.5 * (gq.f2 - 4. * gm.f2 + 3 * gc.f2);
        s3 = s3 + (fp.f3 - fm.f3) + 0.5 * (gq.f3 - 4. * gm.f3 + 3 * gc.f3);
        s4 = s4 + (fp.f4 - fm.f4) + 0.5 * (gq.f4 - 4. * gm.f4 + 3 * gc.f4);
      }
    }
    M.s[1][i] = s2;
    M.s[2][i] = s3;
    M.s[3][i] = s4;
    sv[0] = M.s[0][i]; sv[1] = s2; sv[2] = s3; sv[3] = s4;
  }
  if (norms) {
    const ResTile T{(int)blockIdx.y, j0, k0, (int)blockIdx.x};
    norm_tile(M, T, j, k, valid, sv, &red[0][0]);
  }
}

// ========================================================================= SA
// constants of src/sadata.h:2-4
constexpr double SA_cb1 = 0.1355, SA_sigma = 0.6666667, SA_cb2 = 0.622, SA_akt = 0.41,
                 SA_cw1 = 3.2390678, SA_cw2 = 0.3, SA_cw3 = 2., SA_cv1 = 7.1, SA_ct3 = 1.2, SA_ct4 = 0.5;

// integer powers as the reference build evaluates them (left to right)
__device__ __forceinline__ double ip3(double x) { return (x * x) * x; }
__device__ __forceinline__ double ip6(double x) { return ((((x * x) * x) * x) * x) * x; }

// laminar viscosity (lamvis, src/samodel2d.f:200-238) and the velocities
// vmu_sa keeps (src/samodel2d.f:61-72), all cells
// ovl: the state as ilu2d's primitive round trip leaves it
__global__ __launch_bounds__(256) void k_sa_prim(const DevMesh* __restrict__ Ms, int ovl) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  long long n2 = (long long)M.jd * M.kd;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    double q1 = M.q[0][i], q2 = M.q[1][i], q3 = M.q[2][i], q4 = M.q[3][i];
    if (ovl) q_roundtrip(q1, q2, q3, q4, M.q[M.nq - 1][i]);
    sa_prim_cell(M, i, q1, q2, q3, q4);
  }
}

__device__ __forceinline__ int sa_periodic_src(int idir, int js, int je, int ks, int ke, int jd, int kd, int jk) {
  if (idir == 1) { int jj = je - js + 1; return jd - 2 * jj + (jk - js + 1); }
  if (idir == -1) { int jj = je - js + 1; return 1 + 2 * jj - (je - jk + 1); }
  if (idir == 2) { int kk = ke - ks + 1; return kd - 2 * kk + (jk - ks + 1); }
  int kk = ke - ks + 1;
  return 1 + 2 * kk - (ke - jk + 1);
}

// c_turm (src/samodel2d.f:372-399) for one cell, from its q(1), q(nq),
// laminar viscosity and q(5)
__device__ __forceinline__ double c_turm_val(double q1, double qn, double vmul, double q5) {
  double rj = q1 * qn;
  double vnul = vmul / rj;
  double chi = q5 / vnul;
  double fv1 = ip3(chi) / (ip3(chi) + ip3(SA_cv1));
  return q1 * qn * q5 * fv1;
}
// the eddy viscosity of cell i; with ovl the SA model runs beside ilu2d, so
// q1 is taken through ilu2d's round trip, (q1 qj) / qj, and the value goes
// to turmu_n
__device__ __forceinline__ void put_turm(const DevMesh& M, long long i, double q1, double qj, double vmul, double q5,
                                         int ovl) {
  if (ovl) M.turmu_n[i] = c_turm_val((q1 * qj) / qj, qj, vmul, q5);
  else M.turmu[i] = c_turm_val(q1, qj, vmul, q5);
}

__device__ __forceinline__ void c_turm_cell(const DevMesh& M, long long i, int ovl) {
  put_turm(M, i, M.q[0][i], M.q[M.nq - 1][i], M.sa_vmul[i], M.q[4][i], ovl);
}

// turbc (src/samodel2d.f:827-1303) for one boundary segment
__device__ void turbc_segment(const DevMesh& M, const DevBC& b) {
  double* q5 = M.q[4];
  const double* u = M.sa_u;
  const double* v = M.sa_v;
  const int t = threadIdx.x, nt = blockDim.x;
  {
    int js = b.js, je = b.je, ks = b.ks, ke = b.ke, idir = b.idir;
    int iadd = idir >= 0 ? 1 : -1, iadir = idir >= 0 ? idir : -idir;
    int ty = b.type;
    double foso = 1.0;
    if (ty == 3 || ty == 4) {  // turbc_extpt, one halo layer
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd, j2 = j1 + iadd;
        for (int k = ks + t; k <= ke; k += nt)
          q5[cell(M, j, k)] = (1. + foso) * q5[cell(M, j1, k)] - foso * q5[cell(M, j2, k)];
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd, k2 = k1 + iadd;
        for (int j = js + t; j <= je; j += nt)
          q5[cell(M, j, k)] = (1. + foso) * q5[cell(M, j, k1)] - foso * q5[cell(M, j, k2)];
      }
    } else if (ty == 5) {  // turbc_wall
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd;
        for (int k = ks + t; k <= ke; k += nt) q5[cell(M, j, k)] = -q5[cell(M, j1, k)];
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd;
        for (int j = js + t; j <= je; j += nt) q5[cell(M, j, k)] = -q5[cell(M, j, k1)];
      }
    } else if (ty == 10 || ty == 11) {  // turbc_sym
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = (idir == 1) ? je + 1 : js - 1;
        for (int k = ks + t; k <= ke; k += nt) q5[cell(M, j, k)] = q5[cell(M, j1, k)];
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
        for (int j = js + t; j <= je; j += nt) q5[cell(M, j, k)] = q5[cell(M, j, k1)];
      }
    } else if (ty == 22) {  // turbc_periodic
      int nj = je - js + 1, nk = ke - ks + 1;
      for (int w = t; w < nj * nk; w += nt) {
        int j = js + w % nj, k = ks + w / nj;
        long long src = (iadir == 1) ? cell(M, sa_periodic_src(idir, js, je, ks, ke, M.jd, M.kd, j), k)
                                     : cell(M, j, sa_periodic_src(idir, js, je, ks, ke, M.jd, M.kd, k));
        q5[cell(M, j, k)] = q5[src];
      }
    } else if (ty == 51 && iadir == 2) {  // turbc_wake
      int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
      for (int j = js + t; j <= je; j += nt) {
        int jj = M.jd - j + 1;
        double a = q5[cell(M, jj, k1)], c = q5[cell(M, j, k1)];
        q5[cell(M, j, k)] = a;
        q5[cell(M, jj, k)] = c;
      }
    } else if (ty == 46) {  // turbc_inf
      int nj = je - js + 1, nk = ke - ks + 1;
      for (int w = t; w < nj * nk; w += nt) q5[cell(M, js + w % nj, ks + w / nj)] = M.c.vnuinf;
    } else if (ty == 47) {  // turbc_out
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j1, k);
          double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
          double ug1 = M.has_gv ? M.ug[i1] : 0.0, vg1 = M.has_gv ? M.vg[i1] : 0.0;
          double val = M.c.vnuinf;
          double uu = (u[i] - ug + u[i1] - ug1) * M.xx[i];
          uu = uu + (v[i] - vg + v[i1] - vg1) * M.xy[i];
          uu = uu * iadd;
          if (uu < 0.) val = q5[i1];
          q5[i] = val;
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j, k1);
          double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
          double ug1 = M.has_gv ? M.ug[i1] : 0.0, vg1 = M.has_gv ? M.vg[i1] : 0.0;
          double val = M.c.vnuinf;
          double uu = (u[i] - ug + u[i1] - ug1) * M.yx[i];
          uu = uu + (v[i] - vg + v[i1] - vg1) * M.yy[i];
          uu = uu * iadd;
          if (uu < 0.) val = q5[i1];
          q5[i] = val;
        }
      }
    }
  }
}

// the segments of one phase of every instance, one workgroup per segment
// (segments of a phase touch disjoint cells; phases run in the reference's
// segment order)
__global__ __launch_bounds__(BC_THREADS) void k_turbc(const DevMesh* __restrict__ Ms, int phase) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int ib = blockIdx.x;
  if (ib >= M.nbc || M.bc[ib].phase != phase) return;
  turbc_segment(M, M.bc[ib]);
}

// the final turbc and the eddy viscosity of the halo ring after it
// (c_turm), one workgroup per instance: every segment in order (one launch
// here costs less than a launch per phase and one for the ring)
__global__ __launch_bounds__(BC_THREADS) void k_turbc_eddy(const DevMesh* __restrict__ Ms, int ovl) {
  const DevMesh& M = Ms[blockIdx.x];
  if (*M.status) return;
  for (int ib = 0; ib < M.nbc; ++ib) {
    turbc_segment(M, M.bc[ib]);
    __syncthreads();
  }
  const int jd = M.jd, kd = M.kd;
  const int nring = 2 * jd + 2 * (kd - 2);
  for (int w = threadIdx.x; w < nring; w += blockDim.x) {
    int j, k;
    if (w < jd) { j = w + 1; k = 1; }
    else if (w < 2 * jd) { j = w - jd + 1; k = kd; }
    else if (w < 2 * jd + kd - 2) { j = 1; k = w - 2 * jd + 2; }
    else { j = jd; k = w - 2 * jd - (kd - 2) + 2; }
    c_turm_cell(M, cell(M, j, k), ovl);
  }
}

// convective upwind speeds of a cell (src/samodel2d.f:483-491)
struct UpDown { double up, um, vp, vm; };
__device__ __forceinline__ UpDown sa_updown(const DevMesh& M, long long i) {
  double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
  double u = M.sa_u[i], v = M.sa_v[i];
  double uu = M.xx[i] * (u - ug) + M.xy[i] * (v - vg);
  double vv = M.yx[i] * (u - ug) + M.yy[i] * (v - vg);
  UpDown o;
  o.up = 0.5 * (uu + fabs(uu)); o.um = 0.5 * (uu - fabs(uu));
  o.vp = 0.5 * (vv + fabs(vv)); o.vm = 0.5 * (vv - fabs(vv));
  return o;
}

__device__ __forceinline__ double sa_vnul(const DevMesh& M, long long i) {
  return M.sa_vmul[i] / M.q[0][i] / M.q[M.nq - 1][i];
}

// diffusion half-point coefficient chp and difference dnuhp of the face
// between cells a and b (src/samodel2d.f:575-584, 622-631)
__device__ __forceinline__ void sa_half(const DevMesh& M, long long a, long long b, double rey, double& chp,
                                        double& dn) {
  double vnulh = 0.5 * (sa_vnul(M, a) + sa_vnul(M, b));
  double vnuh = 0.5 * (M.q[4][a] + M.q[4][b]);
  chp = (1. + SA_cb2) / SA_sigma / rey * (vnulh + vnuh);
  dn = M.q[4][b] - M.q[4][a];
}

// rhslhs (convec + diffus + source) and the pseudo-time scaling of vmu_sa
// (src/samodel2d.f:101-131, 402-824) for every interior cell, in gather
// form: each cell assembles its own row of the two tridiagonal systems.
__global__ __launch_bounds__(256) void k_sa_rhs(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const DevConst& C = M.c;
  const int js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const int nj = je - js + 1, nk = ke - ks + 1;
  const double rey = C.rey;
  const double* q5 = M.q[4];
  const double* xx = M.xx;
  const double* xy = M.xy;
  const double* yx = M.yx;
  const double* yy = M.yy;
  long long ncell = (long long)nj * nk;
  for (long long w = blockIdx.x * 256ll + threadIdx.x; w < ncell; w += 256ll * gridDim.x) {
    const int j = js + (int)(w % nj), k = ks + (int)(w / nj);
    const long long i = cell(M, j, k);
    const long long iw = cell(M, j - 1, k), ie = cell(M, j + 1, k), is = cell(M, j, k - 1), in = cell(M, j, k + 1);
    UpDown c = sa_updown(M, i);
    UpDown cw = sa_updown(M, iw), ce = sa_updown(M, ie), cs = sa_updown(M, is), cn = sa_updown(M, in);
    double nt = q5[i];
    // ---- convec (src/samodel2d.f:493-538)
    double sj = M.s[4][i];
    sj = sj - c.up * (nt - q5[iw]) - c.um * (q5[ie] - nt);
    sj = sj - c.vp * (nt - q5[is]) - c.vm * (q5[in] - nt);
    double aj = 0.0, bj = 0.0, cj = 0.0, ak = 0.0, bk = 0.0, ck = 0.0;
    if (j > js) aj = aj - (c.up > 1.0e-12 ? 1.0 : 0.0) * (cw.up + cw.um);
    bj = bj + c.up - c.um;
    if (j < je) cj = cj + (c.um < -1.0e-12 ? 1.0 : 0.0) * (ce.up + ce.um);
    if (k > ks) ak = ak - (c.vp > 1.0e-12 ? 1.0 : 0.0) * (cs.vp + cs.vm);
    bk = bk + c.vp - c.vm;
    if (k < ke) ck = ck + (c.vm < -1.0e-12 ? 1.0 : 0.0) * (cn.vp + cn.vm);
    // ---- diffus, j direction (src/samodel2d.f:570-607)
    {
      double chpm, dnm, chpp, dnp;
      sa_half(M, iw, i, rey, chpm, dnm);
      sa_half(M, i, ie, rey, chpp, dnp);
      double dmxhm = 0.5 * (xx[iw] + xx[i]), dmyhm = 0.5 * (xy[iw] + xy[i]);
      double dmxhp = 0.5 * (xx[i] + xx[ie]), dmyhp = 0.5 * (xy[i] + xy[ie]);
      double dxp = dmxhp * xx[i] + dmyhp * xy[i];
      double dxm = dmxhm * xx[i] + dmyhm * xy[i];
      double c2 = SA_cb2 / SA_sigma / rey;
      c2 = c2 * (sa_vnul(M, i) + nt);
      double dcp = dxp * (chpp - c2);
      double dcm = dxm * (chpm - c2);
      double ax = fmax(dcm, 0.0), cx = fmax(dcp, 0.0);
      sj = sj - ax * dnm + cx * dnp;
      aj = aj - ax;
      cj = cj - cx;
      bj = bj + ax + cx;
    }
    // ---- diffus, k direction (src/samodel2d.f:617-654)
    {
      double chpm, dnm, chpp, dnp;
      sa_half(M, is, i, rey, chpm, dnm);
      sa_half(M, i, in, rey, chpp, dnp);
      double dmxhm = 0.5 * (yx[is] + yx[i]), dmyhm = 0.5 * (yy[is] + yy[i]);
      double dmxhp = 0.5 * (yx[i] + yx[in]), dmyhp = 0.5 * (yy[i] + yy[in]);
      double dyp = dmxhp * yx[i] + dmyhp * yy[i];
      double dym = dmxhm * yx[i] + dmyhm * yy[i];
      double c2 = SA_cb2 / SA_sigma / rey;
      c2 = c2 * (sa_vnul(M, i) + nt);
      double dcp = dyp * (chpp - c2);
      double dcm = dym * (chpm - c2);
      double ay = fmax(dcm, 0.0), cy = fmax(dcp, 0.0);
      sj = sj - ay * dnm + cy * dnp;
      ak = ak - ay;
      ck = ck - cy;
      bk = bk + ay + cy;
    }
    // ---- source (src/samodel2d.f:729-811), isour = 0, fturf = 0
    {
      double vort = 0.0;
      if (M.bodyflag) {
        double ux = (M.sa_u[ie] - M.sa_u[iw]) * 0.5;
        double vx = (M.sa_v[ie] - M.sa_v[iw]) * 0.5;
        double uy = (M.sa_u[in] - M.sa_u[is]) * 0.5;
        double vy = (M.sa_v[in] - M.sa_v[is]) * 0.5;
        vort = fabs(xy[i] * ux - xx[i] * vx + yy[i] * uy - yx[i] * vy);
      }
      const double d2min = 1.e-12, rmax = 10.0, stilim = 1.e-12, fturf = 0.0;
      const double cv1_3 = ip3(SA_cv1), cw3_6 = ip6(SA_cw3);
      double vnul = sa_vnul(M, i);
      double chi = nt / vnul;
      double d = M.sa_sn[i];
      chi = fmax(chi, 1.e-12);
      double chi3 = ip3(chi);
      double fv1 = chi3 / (chi3 + cv1_3);
      double fv2 = 1. - chi / (1. + chi * fv1);
      double fv3 = 1.0;
      double ft2 = fturf * SA_ct3 * exp(-1. * SA_ct4 * chi * chi);
      double dchi = 1. / vnul;
      double dfv1 = (3. * cv1_3) * (chi * chi) * dchi * sqd(1. / (chi3 + cv1_3));
      double dfv2 = (fv2 - 1.) / chi / vnul + (1. - fv2) * (1 - fv2) * (dfv1 + fv1 / chi / vnul);
      double dfv3 = 0.0;
      double dft2 = (-2. * SA_ct4) * chi * dchi * ft2;
      double d2 = fmax(d * d, d2min);
      double stilda = vort * fv3 + vnul / (d2 * SA_akt * SA_akt * rey) * chi * fv2;
      stilda = fmax(stilda, stilim);
      double r = nt / (d2 * SA_akt * SA_akt * rey) / stilda;
      r = fmin(r, rmax);
      double g = r + SA_cw2 * (ip6(r) - r);
      double g6 = ip6(g);
      double fw = (1. + cw3_6) / (g6 + cw3_6);
      fw = g * pow(fw, 1. / 6.);
      double dstild = vort * dfv3 + vnul * (dchi * fv2 + chi * dfv2) / (d2 * SA_akt * SA_akt * rey);
      double dr = vnul * (dchi - chi * dstild / stilda) / stilda / (d2 * SA_akt * SA_akt * rey);
      double dg = dr * (1. + SA_cw2 * (6. * pow(r, 5.) - 1.));
      double dfw = pow((1. + cw3_6) / (g6 + cw3_6), 1. / 6.);
      dfw = dfw * dg * (1. - g6 / (g6 + cw3_6));
      double pro = SA_cb1 * stilda * (1. - ft2);
      double des = (SA_cw1 * fw - SA_cb1 / (SA_akt * SA_akt) * ft2) / d2 / rey;
      double prod = SA_cb1 * stilda * (1. - ft2) * nt;
      double dest = (SA_cw1 * fw - SA_cb1 / (SA_akt * SA_akt) * ft2) * nt * nt / d2 / rey;
      double dpro = pro * dstild / stilda - SA_cb1 * stilda * dft2;
      double ddes = (SA_cw1 * dfw - SA_cb1 / (SA_akt * SA_akt) * dft2) / d2 / rey * vnul * chi;
      ddes = ddes + des;
      ddes = ddes * nt;
      dpro = dpro * nt;
      sj = sj + prod - dest;
      double tk1 = fmax(des * nt - pro, 0.0);
      double tk2 = fmax(ddes - dpro, 0.0);
      bk = bk + tk1 + tk2;
    }
    // ---- scale by the turbulence pseudo time step
    double tscal = M.tscale[i];
    if (C.timeac == 1.) {
      double qj = M.q[M.nq - 1][i];
      tscal = (double)imx(M.iblank[i], 0) * (1.0 + 0.002 * sqrt(qj)) / (1. + sqrt(qj));
      tscal = tscal * 1.0;
      tscal = tscal / (1. + tscal / C.h);
    }
    M.sa_sj[i] = sj * tscal;
    M.sa_aj[i] = aj * tscal;
    M.sa_cj[i] = cj * tscal;
    M.sa_ak[i] = ak * tscal;
    M.sa_ck[i] = ck * tscal;
    M.sa_bb[i] = 1. + (bj + bk) * tscal;
  }
}

// The two scalar tridiagonal solves of vmu_sa are sequential recurrences
// along each grid line, both eliminated from the line's two ends at once
// (thomas2_lds), which halves the dependent chain; lines of up to 32 cells
// keep the reference's one-sided recurrence (lsolvej / lsolvek).  One wave
// owns one line: all lanes stage the line's coefficients in LDS (stage_line4:
// every load in flight at once, coalesced along j, gathered along k), two
// lanes run the two sides out of LDS (reads a chunk ahead of the dependent
// chain), and all lanes write the results.  The wave, tile and lane kernels
// of a direction evaluate the same operations in the same order, so every
// mapping gives bit-identical results; against the one-sided recurrence the
// results differ in rounding only.
// lsolvej (src/samodel2d.f:1306-1344), eliminated from both ends, and s1,
// sjmat*bjmat of vmu_sa (src/samodel2d.f:148-152) for the k row blockIdx.x
__global__ __launch_bounds__(64) void k_sa_solve_j(const DevMesh* __restrict__ Ms) {
  __shared__ double la[SA_LINE_MAX], lb[SA_LINE_MAX], lc[SA_LINE_MAX], ls[SA_LINE_MAX];
  __shared__ double lcp[SA_LINE_MAX], lsp[SA_LINE_MAX];
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int js = M.js, je = M.je;
  const int k = M.ks + blockIdx.x;
  if (k > M.ke) return;
  const int n = je - js + 1;
  const long long i0 = cell(M, js, k);
  double* __restrict__ s = M.sa_sj;
  const double* __restrict__ c = M.sa_cj;
  const double* __restrict__ a = M.sa_aj;
  const double* __restrict__ b = M.sa_bb;
  stage_line4(a, b, c, s, i0, 1, n, la, lb, lc, ls);
  __syncthreads();
  int nr, m;
  sa_split(n, nr, m);
  if (threadIdx.x < (nr > 0 ? 2u : 1u)) thomas2_lds<1, 1>(la, lb, lc, ls, lsp, lcp, n, (int)threadIdx.x);
  __syncthreads();
  // sjmat = (s1 - s2) * bjmat with s2 = 0
  for (int t = threadIdx.x; t < n; t += 64) s[i0 + t] = (0. + lsp[t] - 0.) * lb[t];
}

// lsolvek (src/samodel2d.f:1347-1385), s2 and the nu-tilde update with its
// clip (src/samodel2d.f:157-188), then the interior eddy viscosity (c_turm)
// for the j column blockIdx.x
__global__ __launch_bounds__(64) void k_sa_solve_k(const DevMesh* __restrict__ Ms, int ovl) {
  __shared__ double la[SA_LINE_MAX], lb[SA_LINE_MAX], lc[SA_LINE_MAX], ls[SA_LINE_MAX];
  __shared__ double lcp[SA_LINE_MAX], lsp[SA_LINE_MAX];
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int ks = M.ks, ke = M.ke;
  const int j = M.js + blockIdx.x;
  if (j > M.je) return;
  const int n = ke - ks + 1;
  const long long i0 = cell(M, j, ks);
  const long long st = M.jd;
  const double* __restrict__ s = M.sa_sj;
  const double* __restrict__ c = M.sa_ck;
  const double* __restrict__ a = M.sa_ak;
  const double* __restrict__ b = M.sa_bb;
  stage_line4(a, b, c, s, i0, st, n, la, lb, lc, ls);
  __syncthreads();
  int nr, m;
  sa_split(n, nr, m);
  if (threadIdx.x < (nr > 0 ? 2u : 1u)) thomas2_lds<1, 1>(la, lb, lc, ls, lsp, lcp, n, (int)threadIdx.x);
  __syncthreads();
  const double vnulim = 1.0e-20, relfac = 1.0;
  for (int t = threadIdx.x; t < n; t += 64) {
    long long i = i0 + t * st;
    double dq = relfac * (0. + lsp[t]);
    double q5 = M.q[4][i];
    q5 = fmax((q5 + dq * (double)imx(M.iblank[i], 0)), vnulim);
    M.q[4][i] = q5;
    c_turm_cell(M, i, ovl);
  }
}

// lsolvek, the nu-tilde update and c_turm for TJ adjacent j columns
// (blockIdx.x) at once, for lines of up to nmax cells: the workgroup (32*TJ
// threads) stages the TJ k lines as rows of TJ consecutive cells (each row one
// TJ*8-byte run, where the one-line kernel gathers one element per cache
// line; TJ = 16 makes each run a whole 128-byte line), lanes 0..TJ-1 of wave
// 0 run the TJ recurrences out of LDS (column c at stride TJ), and all
// threads apply the update and the eddy viscosity, a few cells each with
// their loads issued together.  Columns past je stage a harmless b = 1 row
// and are never written.  LDS: five arrays of nmax*TJ doubles (dynamic).
constexpr int SAK_TJ = 8, SAK_NMAX = 256;
constexpr int SAK_TJ16 = 16, SAK_NMAX16 = 192;   // 16 columns for lines of up to 192 cells
template <int TJ>
__global__ __launch_bounds__(32 * TJ) void k_sa_solve_kt(const DevMesh* __restrict__ Ms, int ovl, int nmax) {
  extern __shared__ __attribute__((aligned(16))) double sak_lds[];
  double* const la = sak_lds;
  double* const lb = la + (size_t)nmax * TJ;
  double* const lc = lb + (size_t)nmax * TJ;
  double* const ls = lc + (size_t)nmax * TJ;
  double* const lcp = ls + (size_t)nmax * TJ;
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int ks = M.ks, ke = M.ke;
  const int j0 = M.js + (int)blockIdx.x * TJ;
  if (j0 > M.je) return;
  const int n = ke - ks + 1;
  const int ne = n * TJ;
  constexpr int NT = 32 * TJ;   // threads
  const int tid = (int)threadIdx.x;
  const double* __restrict__ s = M.sa_sj;
  const double* __restrict__ c = M.sa_ck;
  const double* __restrict__ a = M.sa_ak;
  const double* __restrict__ b = M.sa_bb;
  // element e = k*TJ + col of the tile; 4 elements per thread in flight at once
  constexpr int U = 4;
  for (int e0 = 0; e0 < ne; e0 += NT * U) {
    double ra[U], rb[U], rc[U], rs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + tid + NT * u;
      const int j = j0 + (e & (TJ - 1));
      const int k = ks + (e < ne ? e : ne - 1) / TJ;
      const long long i = cell(M, j <= M.je ? j : M.je, k);
      ra[u] = a[i]; rb[u] = b[i]; rc[u] = c[i]; rs[u] = s[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + tid + NT * u;
      if (e < ne) {
        const bool in = j0 + (e & (TJ - 1)) <= M.je;
        la[e] = in ? ra[u] : 0.0; lb[e] = in ? rb[u] : 1.0; lc[e] = in ? rc[u] : 0.0; ls[e] = in ? rs[u] : 0.0;
      }
    }
  }
  __syncthreads();
  {
    // column tid & 31 of the tile, side tid >> 5 (lanes 0..TJ-1 and 32..32+TJ-1 of wave 0)
    int nr, m;
    sa_split(n, nr, m);
    const int cl = tid & 31, sd = tid >> 5;
    if (cl < TJ && (sd == 0 || (sd == 1 && nr > 0)))
      thomas2_lds<TJ, 32>(la + cl, lb + cl, lc + cl, ls + cl, ls + cl, lcp + cl, n, sd);
  }
  __syncthreads();
  // the update and c_turm, U cells per thread with all their loads issued
  // before the first store
  const double vnulim = 1.0e-20, relfac = 1.0;
  for (int e0 = 0; e0 < ne; e0 += NT * U) {
    long long ii[U];
    bool ok[U];
    double q5[U], q1[U], qj[U], vm[U], ib[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + tid + NT * u;
      const int j = j0 + (e & (TJ - 1));
      ok[u] = e < ne && j <= M.je;
      ii[u] = cell(M, j <= M.je ? j : M.je, ks + (e < ne ? e : ne - 1) / TJ);
      q5[u] = M.q[4][ii[u]]; q1[u] = M.q[0][ii[u]]; qj[u] = M.q[M.nq - 1][ii[u]]; vm[u] = M.sa_vmul[ii[u]];
      ib[u] = (double)imx(M.iblank[ii[u]], 0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      const int e = e0 + tid + NT * u;
      double dq = relfac * (0. + ls[e]);
      const double v = fmax((q5[u] + dq * ib[u]), vnulim);
      M.q[4][ii[u]] = v;
      put_turm(M, ii[u], q1[u], qj[u], vm[u], v, ovl);
    }
  }
}

// Line solves for many lines at once: one lane per line (along j: one lane
// per side of a line), the same arithmetic as the one-wave kernels.
// k_sa_lines_j: lines along j (one per k row), 32 rows per workgroup, lane
// l = 32h + r of wave 0 running side h of row r's two-ended elimination
// (h = 0 the left part from j = js, h = 1 the right part from j = je).
// Each side is cut into 32-cell tiles, counted from its own end; tile x of
// side 1 is stored reversed (position p holds cell n-1-32x-p), so both
// sides walk positions upwards in the elimination and downwards in the
// back substitution.  Waves 1..SAJ_MW move the tiles (row-contiguous
// segments) between global memory and two LDS slots: while wave 0 works on
// tile x in one slot, the movers write tile x-1's results back from the
// other slot and refill it with tile x+1, whose loads they issued SAJ_D
// tiles earlier, so the recurrences do not wait on memory.  The elimination
// leaves s' and c' (a' on the right) in place, the back substitution writes
// s = (s1 - s2) * bjmat.
constexpr int SA_TJ = SAJ_TILE, SAJ_MW = 4, SAJ_D = 2;
constexpr int SAJ_THREADS = 64 * (1 + SAJ_MW);
constexpr int SAJ_PER = 64 * SA_TJ / (64 * SAJ_MW);   // elements per mover lane and array
// total lines (both directions, all instances) above which the solves run one lane per line
constexpr long long SA_LANE_LINES = 4096;

__global__ __launch_bounds__(SAJ_THREADS) void k_sa_lines_j(const DevMesh* __restrict__ Ms) {
  __shared__ double ta[2][64][SA_TJ + 1], tb[2][64][SA_TJ + 1], tc[2][64][SA_TJ + 1], ts[2][64][SA_TJ + 1];
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const int k0 = ks + 32 * blockIdx.x;
  if (k0 > ke) return;
  const int nrow = ke - k0 + 1 < 32 ? ke - k0 + 1 : 32;
  const int n = je - js + 1;
  int nr, m;
  sa_split(n, nr, m);
  const int nl = (m + SA_TJ - 1) / SA_TJ;   // tiles of the left side (nl >= nr)
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, ml = (int)threadIdx.x - 64;
  const int h = l >> 5, rr = l & 31;         // the solver lane's side and row
  // tiles of side h, and the last position of its last tile
  const int ntl = h ? nr : nl;
  const int lastp = h ? SA_TJ - 1 : m - 1 - SA_TJ * (nl - 1);
  gdouble* __restrict__ sS = gptr(M.sa_sj);
  gdouble* __restrict__ sC = gptr(M.sa_cj);
  const gdouble* __restrict__ sA = gptr(M.sa_aj);
  const gdouble* __restrict__ sB = gptr(M.sa_bb);
  // element p of mover lane ml in a tile: e = ml + 256 p, slot row e / SA_TJ
  // (side and row as for the solver lanes), position e % SA_TJ; the global
  // index is clamped into the grid, so every load is unconditional and no
  // loaded value merges across a branch
  auto gidx = [&](int x, int p, int& r, int& cc, bool& ok) {
    const int e = ml + 64 * SAJ_MW * p;
    r = e / SA_TJ; cc = e % SA_TJ;
    const int hh = r >> 5, rw = r & 31;
    const int t = hh ? n - 1 - SA_TJ * x - cc : SA_TJ * x + cc;   // cell of the line
    ok = rw < nrow && (hh ? x < nr : t < m);
    const int tc = t < 0 ? 0 : (t < n ? t : n - 1);
    return cell(M, js + tc, k0 + (rw < nrow ? rw : 0));
  };
  double ra[SAJ_D][SAJ_PER], rb[SAJ_D][SAJ_PER], rc[SAJ_D][SAJ_PER], rs[SAJ_D][SAJ_PER];

  // ---- elimination from both ends
  auto fetch_f = [&](int t, double (&a)[SAJ_PER], double (&b)[SAJ_PER], double (&c)[SAJ_PER], double (&s)[SAJ_PER]) {
#pragma unroll
    for (int p = 0; p < SAJ_PER; ++p) {
      int r, cc; bool ok;
      const long long i = gidx(t, p, r, cc, ok);
      (void)ok;
      a[p] = sA[i]; b[p] = sB[i]; c[p] = sC[i]; s[p] = sS[i];
    }
  };
  // elements outside a side (rows past the last, cells past the side's end
  // in its last tile) hold copies of clamped cells: the recurrences never
  // read them
  auto put_f = [&](int slot, const double (&a)[SAJ_PER], const double (&b)[SAJ_PER], const double (&c)[SAJ_PER],
                   const double (&s)[SAJ_PER]) {
#pragma unroll
    for (int p = 0; p < SAJ_PER; ++p) {
      const int e = ml + 64 * SAJ_MW * p, r = e / SA_TJ, cc = e % SA_TJ;
      ta[slot][r][cc] = a[p]; tb[slot][r][cc] = b[p]; tc[slot][r][cc] = c[p]; ts[slot][r][cc] = s[p];
    }
  };
  auto back_f = [&](int t, int slot) {   // s' and c' (a') of tile t to global memory
#pragma unroll
    for (int p = 0; p < SAJ_PER; ++p) {
      int r, cc; bool ok;
      const long long i = gidx(t, p, r, cc, ok);
      if (ok) { sS[i] = ts[slot][r][cc]; sC[i] = tc[slot][r][cc]; }
    }
  };
  double cp = 0.0, sp = 0.0;
  // side h's lower and upper coefficients: (a, c) on the left, (c, a) on the
  // right; c' / a' goes where c was
  auto solve_f = [&](int x, int slot) {
    if (rr >= nrow || x >= ntl) return;
    const int cn = x == ntl - 1 ? lastp + 1 : SA_TJ;   // positions of this tile
    double (*LO)[SA_TJ + 1] = h ? tc[slot] : ta[slot];
    double (*HI)[SA_TJ + 1] = h ? ta[slot] : tc[slot];
    double (*B)[SA_TJ + 1] = tb[slot];
    double (*Cc)[SA_TJ + 1] = tc[slot];
    double (*S)[SA_TJ + 1] = ts[slot];
    auto elem = [&](int cc) {
      double ai = LO[l][cc];
      double bb = 1. / (B[l][cc] - ai * cp);
      sp = (S[l][cc] - ai * sp) * bb;
      cp = HI[l][cc] * bb;
      S[l][cc] = sp; Cc[l][cc] = cp;
    };
    if (x == 0) {
      double bb = 1. / B[l][0];
      sp = S[l][0] * bb; cp = HI[l][0] * bb;
      S[l][0] = sp; Cc[l][0] = cp;
#pragma unroll
      for (int cc = 1; cc < SA_TJ; ++cc)
        if (cc < cn) elem(cc);
    } else if (cn == SA_TJ) {
      // chunks of THOMAS_CH cells, each chunk's LDS reads issued before the
      // previous chunk's recurrence
      constexpr int CH = THOMAS_CH;
      double xa[2][CH], xb[2][CH], xc[2][CH], xs[2][CH];
      auto ld = [&](int q, int c0) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          xa[q][u] = LO[l][c0 + u]; xb[q][u] = B[l][c0 + u]; xc[q][u] = HI[l][c0 + u]; xs[q][u] = S[l][c0 + u];
        }
      };
      ld(0, 0);
#pragma unroll
      for (int k = 0; k < SA_TJ / CH; ++k) {
        if (k + 1 < SA_TJ / CH) ld((k + 1) & 1, (k + 1) * CH);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int cc = k * CH + u;
          const double ai = xa[k & 1][u];
          double bb = 1. / (xb[k & 1][u] - ai * cp);
          sp = (xs[k & 1][u] - ai * sp) * bb;
          cp = xc[k & 1][u] * bb;
          S[l][cc] = sp; Cc[l][cc] = cp;
        }
      }
    } else {
#pragma unroll
      for (int cc = 0; cc < SA_TJ; ++cc)
        if (cc < cn) elem(cc);
    }
  };
  // Tile x sits in register set x % SAJ_D.  Every fetch is unconditional
  // (tile indices past the sides are clamped to the last tile) and the
  // movers' main loop runs whole groups of SAJ_D tiles in code of their own,
  // so no loaded value merges across a branch and each load is waited for
  // only where it is stored to LDS.  Solver and movers pass the same
  // sequence of barriers: one after each pass's prologue, one per tile, and
  // one between the passes.
  auto clampt = [&](int t) { return t < nl ? t : nl - 1; };
  auto step_f = [&](int t, double (&a)[SAJ_PER], double (&b)[SAJ_PER], double (&c)[SAJ_PER],
                    double (&s)[SAJ_PER]) {
    if (t >= 1) back_f(t - 1, (t - 1) & 1);
    if (t + 1 < nl) put_f((t + 1) & 1, a, b, c, s);
    fetch_f(clampt(t + 1 + SAJ_D), a, b, c, s);
    __syncthreads();
  };
  if (wave == 0) {
    __syncthreads();
    for (int t = 0; t < nl; ++t) {
      solve_f(t, t & 1);
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int d = 0; d < SAJ_D; ++d) fetch_f(clampt(d), ra[d], rb[d], rc[d], rs[d]);
    put_f(0, ra[0], rb[0], rc[0], rs[0]);
    fetch_f(clampt(SAJ_D), ra[0], rb[0], rc[0], rs[0]);
    __syncthreads();
    int t0 = 0;
    for (; t0 + SAJ_D <= nl; t0 += SAJ_D) {
#pragma unroll
      for (int d = 0; d < SAJ_D; ++d) {
        const int d1 = (d + 1) % SAJ_D;
        step_f(t0 + d, ra[d1], rb[d1], rc[d1], rs[d1]);
      }
    }
#pragma unroll
    for (int d = 0; d < SAJ_D; ++d) {
      const int d1 = (d + 1) % SAJ_D;
      if (t0 + d < nl) step_f(t0 + d, ra[d1], rb[d1], rc[d1], rs[d1]);
    }
    back_f(nl - 1, (nl - 1) & 1);
  }
  // the meeting cells: each side's value from its own (s', c') and the other
  // side's (lanes l and l ^ 32 hold the two sides of one row)
  double sn = sp;
  if (wave == 0 && nr > 0) {
    const double so = __shfl_xor(sp, 32), co = __shfl_xor(cp, 32);
    sn = (sp - cp * so) / (1. - cp * co);
  }

  // ---- back substitution, tiles in reverse (iteration y takes tile
  // nl-1-y of each side); s = (0 + sn - 0) * bjmat.  Each mover lane
  // reloads the elements it wrote itself, so program order makes them
  // visible.
  auto fetch_b = [&](int t, double (&b)[SAJ_PER], double (&c)[SAJ_PER], double (&s)[SAJ_PER]) {
#pragma unroll
    for (int p = 0; p < SAJ_PER; ++p) {
      int r, cc; bool ok;
      const long long i = gidx(t, p, r, cc, ok);
      (void)ok;
      b[p] = sB[i]; c[p] = sC[i]; s[p] = sS[i];
    }
  };
  auto put_b = [&](int slot, const double (&b)[SAJ_PER], const double (&c)[SAJ_PER], const double (&s)[SAJ_PER]) {
#pragma unroll
    for (int p = 0; p < SAJ_PER; ++p) {
      const int e = ml + 64 * SAJ_MW * p, r = e / SA_TJ, cc = e % SA_TJ;
      tb[slot][r][cc] = b[p]; tc[slot][r][cc] = c[p]; ts[slot][r][cc] = s[p];
    }
  };
  auto back_b = [&](int t, int slot) {
#pragma unroll
    for (int p = 0; p < SAJ_PER; ++p) {
      int r, cc; bool ok;
      const long long i = gidx(t, p, r, cc, ok);
      if (ok) sS[i] = ts[slot][r][cc];
    }
  };
  auto solve_b = [&](int x, int slot) {
    if (rr >= nrow || x >= ntl) return;
    double (*B)[SA_TJ + 1] = tb[slot];
    double (*Cc)[SA_TJ + 1] = tc[slot];
    double (*S)[SA_TJ + 1] = ts[slot];
    auto elem = [&](int cc) {
      sn = S[l][cc] - Cc[l][cc] * sn;
      S[l][cc] = (0. + sn - 0.) * B[l][cc];
    };
    if (x == ntl - 1) {
      // the side's meeting cell starts the substitution
      S[l][lastp] = (0. + sn - 0.) * B[l][lastp];
#pragma unroll
      for (int cc = SA_TJ - 2; cc >= 0; --cc)
        if (cc < lastp) elem(cc);
    } else {
      // chunks of THOMAS_CH cells from the tile's end, reads a chunk ahead
      constexpr int CH = THOMAS_CH;
      double xb[2][CH], xc[2][CH], xs[2][CH];
      auto ld = [&](int q, int c1) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          xb[q][u] = B[l][c1 - u]; xc[q][u] = Cc[l][c1 - u]; xs[q][u] = S[l][c1 - u];
        }
      };
      ld(0, SA_TJ - 1);
#pragma unroll
      for (int k = 0; k < SA_TJ / CH; ++k) {
        if (k + 1 < SA_TJ / CH) ld((k + 1) & 1, SA_TJ - 1 - (k + 1) * CH);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int cc = SA_TJ - 1 - k * CH - u;
          sn = xs[k & 1][u] - xc[k & 1][u] * sn;
          S[l][cc] = (0. + sn - 0.) * xb[k & 1][u];
        }
      }
    }
  };
  __syncthreads();
  // iteration y takes tile nl-1-y; register set y % SAJ_D, tiles past the
  // sides' starts clamped to tile 0
  auto rtile = [&](int y) { return y < nl ? nl - 1 - y : 0; };
  auto step_b = [&](int y, double (&b)[SAJ_PER], double (&c)[SAJ_PER], double (&s)[SAJ_PER]) {
    if (y >= 1) back_b(nl - y, (y - 1) & 1);
    if (y + 1 < nl) put_b((y + 1) & 1, b, c, s);
    fetch_b(rtile(y + 1 + SAJ_D), b, c, s);
    __syncthreads();
  };
  if (wave == 0) {
    __syncthreads();
    for (int y = 0; y < nl; ++y) {
      solve_b(nl - 1 - y, y & 1);
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int d = 0; d < SAJ_D; ++d) fetch_b(rtile(d), rb[d], rc[d], rs[d]);
    put_b(0, rb[0], rc[0], rs[0]);
    fetch_b(rtile(SAJ_D), rb[0], rc[0], rs[0]);
    __syncthreads();
    int y0 = 0;
    for (; y0 + SAJ_D <= nl; y0 += SAJ_D) {
#pragma unroll
      for (int d = 0; d < SAJ_D; ++d) {
        const int d1 = (d + 1) % SAJ_D;
        step_b(y0 + d, rb[d1], rc[d1], rs[d1]);
      }
    }
#pragma unroll
    for (int d = 0; d < SAJ_D; ++d) {
      const int d1 = (d + 1) % SAJ_D;
      if (y0 + d < nl) step_b(y0 + d, rb[d1], rc[d1], rs[d1]);
    }
    back_b(0, (nl - 1) & 1);
  }
}

// k_sa_lines<0>: lines along k (one per j column), one lane per side of a
// line: 32 columns per workgroup, lane l = 32h + c walking side h of column
// c's two-ended elimination (h = 0 up from ks, h = 1 down from ke; the same
// operations as thomas2_lds), with the coefficients of its next SA_PK cells
// loaded ahead into registers.  Lanes on consecutive columns, so every step
// reads one contiguous row segment; the nu-tilde update follows in
// k_sa_update.
constexpr int SA_PK = 16;

template <int JL>
__global__ __launch_bounds__(64) void k_sa_lines(const DevMesh* __restrict__ Ms) {
  static_assert(JL == 0, "lines along j run in k_sa_lines_j");
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  if (js + 32 * (int)blockIdx.x > je) return;
  const int h = (int)threadIdx.x >> 5;
  const int line = js + 32 * blockIdx.x + (threadIdx.x & 31);
  const bool lane_ok = line <= je;
  const int ll = lane_ok ? line : je;
  const int n = ke - ks + 1;
  int nr, m;
  sa_split(n, nr, m);
  // this lane's side: cnt cells, step u at base + u*step (both in elements);
  // a side with no cells (h = 1, nr = 0) walks the left side's cells and
  // stores nothing
  const bool side_ok = h == 0 || nr > 0;
  const bool st_ok = lane_ok && side_ok;
  const int hs = side_ok ? h : 0;
  const int cnt = hs ? n - m : m;
  const long long st = M.jd;
  const long long base = cell(M, ll, ks) + (hs ? (long long)(n - 1) * st : 0);
  const long long step = hs ? -st : st;
  // global (not generic) pointers: the compiler then counts each load's
  // completion exactly and the loads ahead stay in flight
  gdouble* __restrict__ sS = gptr(M.sa_sj);
  gdouble* __restrict__ sC = gptr(M.sa_ck);
  const gdouble* __restrict__ sA = gptr(M.sa_ak);
  const gdouble* __restrict__ sB = gptr(M.sa_bb);
  // lower and upper coefficients of the side: (a, c) going up, (c, a) down;
  // c' / a' goes where c was
  const gdouble* __restrict__ sLo = hs ? (const gdouble*)sC : sA;
  const gdouble* __restrict__ sHi = hs ? sA : (const gdouble*)sC;
  // Loads ahead are unconditional (steps clamped to the side) and the
  // unrolled loops run whole chunks, the remainder separately.  Slot d of a
  // prefetch array holds the step congruent to d modulo SA_PK.
  auto at = [&](int u) { return base + (long long)(u < cnt ? u : cnt - 1) * step; };
  double pa[SA_PK], pb[SA_PK], pc[SA_PK], ps[SA_PK];
#pragma unroll
  for (int d = 0; d < SA_PK; ++d) {
    const long long i = at(d);
    pa[d] = sLo[i]; pb[d] = sB[i]; pc[d] = sHi[i]; ps[d] = sS[i];
  }
  double cp, sp;
  {
    // first cell of the side (its lower coefficient is not used)
    const double b = pb[0], c = pc[0], s = ps[0];
    const long long i = at(SA_PK);
    pa[0] = sLo[i]; pb[0] = sB[i]; pc[0] = sHi[i]; ps[0] = sS[i];
    double bb = 1. / b;
    sp = s * bb; cp = c * bb;
    if (st_ok) { sS[base] = sp; sC[base] = cp; }
  }
  auto fwd = [&](int u, int d) {
    const double a = pa[d], b = pb[d], c = pc[d], s = ps[d];
    const long long i = at(u + SA_PK);
    pa[d] = sLo[i]; pb[d] = sB[i]; pc[d] = sHi[i]; ps[d] = sS[i];
    double bb = 1. / (b - a * cp);
    sp = (s - a * sp) * bb;
    cp = c * bb;
    if (st_ok) { const long long o = base + (long long)u * step; sS[o] = sp; sC[o] = cp; }
  };
  int tb = 1;
  for (; tb + SA_PK <= cnt; tb += SA_PK) {
#pragma unroll
    for (int e = 0; e < SA_PK; ++e) fwd(tb + e, (1 + e) % SA_PK);
  }
#pragma unroll
  for (int e = 0; e < SA_PK; ++e)
    if (tb + e < cnt) fwd(tb + e, (1 + e) % SA_PK);
  // the meeting cells (lanes l and l ^ 32 hold the two sides of a column)
  double sn = sp;
  if (nr > 0) {
    const double so = __shfl_xor(sp, 32), co = __shfl_xor(cp, 32);
    sn = (sp - cp * so) / (1. - cp * co);
  }
  // back substitution towards the side's start (the values this lane just
  // wrote), the next cells loaded ahead; slot d holds back-step v congruent
  // to d (back-step v is step cnt-1-v); dq = s2 goes back into sjmat
  auto atb = [&](int v) { return base + (long long)(v < cnt ? cnt - 1 - v : 0) * step; };
  double qs[SA_PK], qc[SA_PK];
#pragma unroll
  for (int d = 0; d < SA_PK; ++d) {
    const long long i = atb(d);
    qs[d] = sS[i]; qc[d] = sC[i];
  }
  {
    const long long i = atb(SA_PK);
    qs[0] = sS[i]; qc[0] = sC[i];
    if (st_ok) sS[base + (long long)(cnt - 1) * step] = sn;
  }
  auto bwd = [&](int v, int d) {
    const double s = qs[d], c = qc[d];
    const long long i = atb(v + SA_PK);
    qs[d] = sS[i]; qc[d] = sC[i];
    sn = s - c * sn;
    if (st_ok) sS[base + (long long)(cnt - 1 - v) * step] = sn;
  };
  int ub = 1;
  for (; ub + SA_PK <= cnt; ub += SA_PK) {
#pragma unroll
    for (int e = 0; e < SA_PK; ++e) bwd(ub + e, (1 + e) % SA_PK);
  }
#pragma unroll
  for (int e = 0; e < SA_PK; ++e)
    if (ub + e < cnt) bwd(ub + e, (1 + e) % SA_PK);
}

// nu-tilde update with its clip (src/samodel2d.f:157-188) and the interior
// eddy viscosity after the lane-per-line k solve: a pointwise pass over the
// interior cells, so its loads stream at full bandwidth instead of waiting
// on each line's recurrence
__global__ __launch_bounds__(256) void k_sa_update(const DevMesh* __restrict__ Ms, int ovl) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int nj = M.je - M.js + 1, nk = M.ke - M.ks + 1;
  const long long ncell = (long long)nj * nk;
  const double vnulim = 1.0e-20, relfac = 1.0;
  for (long long w = blockIdx.x * 256ll + threadIdx.x; w < ncell; w += 256ll * gridDim.x) {
    const int j = M.js + (int)(w % nj), k = M.ks + (int)(w / nj);
    const long long i = cell(M, j, k);
    const double dq = M.sa_sj[i];
    const double v = fmax((M.q[4][i] + relfac * (0. + dq) * (double)imx(M.iblank[i], 0)), vnulim);
    M.q[4][i] = v;
    put_turm(M, i, M.q[0][i], M.q[M.nq - 1][i], M.sa_vmul[i], v, ovl);
  }
}

// distance to the wall (dist, src/samodel2d.f:243-293); static grids: once
__global__ __launch_bounds__(256) void k_sa_dist(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  const int jd = M.jd, kd = M.kd;
  long long n2 = (long long)jd * kd;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    int j = 1 + (int)(i % jd), k = 1 + (int)(i / jd);
    double sn = 1.e10;
    if (M.bodyflag) {
      if (j >= M.jtail1 && j <= M.jtail2) {
        if (k == 1) sn = 0.0;
        else {
          int jv = j - 1;
          sn = sqrt(sqd(M.x[i] - 0.5 * (M.xv[vert(M, jv, 1)] + M.xv[vert(M, jv + 1, 1)])) +
                    sqd(M.y[i] - 0.5 * (M.yv[vert(M, jv, 1)] + M.yv[vert(M, jv + 1, 1)])));
        }
      } else {
        int jv = M.jtail1 - 1;
        sn = sqrt(sqd(M.x[i] - M.xv[vert(M, jv, 1)]) + sqd(M.y[i] - M.yv[vert(M, jv, 1)]));
      }
    }
    M.sa_sn[i] = sn;
  }
}

static inline int nblk2(long long n, int per) {
  long long b = (n + per - 1) / per;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

static long long max_cells(const DevMesh* hM, int nmi) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  return mx;
}

hipError_t launch_visrhs(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st, int norms) {
  int tiles = 1;
  for (int i = 0; i < nmi; ++i) {
    const int nj = hM[i].je - hM[i].js + 1, nk = hM[i].ke - hM[i].ks + 1;
    const int n = ((nj + VIS_TJ - 1) / VIS_TJ) * ((nk + VIS_TK - 1) / VIS_TK);
    tiles = n > tiles ? n : tiles;
  }
  static bool lds_set = false;
  if (!lds_set) {
    hipError_t e = hipFuncSetAttribute((const void*)k_visrhs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)VIS_LDS);
    if (e != hipSuccess) return e;
    lds_set = true;
  }
  hipLaunchKernelGGL(k_visrhs, dim3(tiles, nmi), dim3(VIS_THREADS), VIS_LDS, st, dM, norms);
  return hipGetLastError();
}

hipError_t launch_sa_dist(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  hipLaunchKernelGGL(k_sa_dist, dim3(nblk2(max_cells(hM, nmi), 256), nmi), dim3(256), 0, st, dM);
  return hipGetLastError();
}

hipError_t launch_sa_prim(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  hipLaunchKernelGGL(k_sa_prim, dim3(nblk2(max_cells(hM, nmi), 256), nmi), dim3(256), 0, st, dM, 0);
  return hipGetLastError();
}

// turbc over every instance: one launch per phase of the segments
static hipError_t launch_turbc(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  int nbc = 1, nph = 1;
  for (int i = 0; i < nmi; ++i) {
    nbc = hM[i].nbc > nbc ? hM[i].nbc : nbc;
    nph = hM[i].nbc_phase > nph ? hM[i].nbc_phase : nph;
  }
  for (int ph = 0; ph < nph; ++ph) {
    hipLaunchKernelGGL(k_turbc, dim3(nbc, nmi), dim3(BC_THREADS), 0, st, dM, ph);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// vmu_sa for every instance (src/samodel2d.f:2-197, itrans = 0)
hipError_t launch_sa(const DevMesh* dM, const DevMesh* hM, int nmi, Marks* mk, hipStream_t st, int sa_lane,
                     int prim_done, int ovl) {
  long long mx = max_cells(hM, nmi);
  int mj = 1, mk_ = 1;
  for (int i = 0; i < nmi; ++i) {
    mj = hM[i].je - hM[i].js + 1 > mj ? hM[i].je - hM[i].js + 1 : mj;
    mk_ = hM[i].ke - hM[i].ks + 1 > mk_ ? hM[i].ke - hM[i].ks + 1 : mk_;
  }
  hipError_t e = hipSuccess;
  if (!prim_done) {
    hipLaunchKernelGGL(k_sa_prim, dim3(nblk2(mx, 256), nmi), dim3(256), 0, st, dM, ovl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  mark(mk, OT2D_ST_SA_PRIM, st);
  if ((e = launch_turbc(dM, hM, nmi, st)) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_BC, st);
  hipLaunchKernelGGL(k_sa_rhs, dim3(nblk2(mx, 256), nmi), dim3(256), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_RHS, st);
  // few lines (one case): one wave per line keeps each recurrence's chain
  // short in latency; many lines (batches): one lane per line
  long long lines = 0;
  for (int i = 0; i < nmi; ++i) lines += (hM[i].ke - hM[i].ks + 1) + (hM[i].je - hM[i].js + 1);
  const bool per_lane = sa_lane < 0 ? lines > SA_LANE_LINES : sa_lane == 1;
  if (per_lane) hipLaunchKernelGGL(k_sa_lines_j, dim3((mk_ + 31) / 32, nmi), dim3(SAJ_THREADS), 0, st, dM);
  else hipLaunchKernelGGL(k_sa_solve_j, dim3(mk_, nmi), dim3(64), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_SOLVE_J, st);
  if (per_lane) {
    hipLaunchKernelGGL(k_sa_lines<0>, dim3((mj + 31) / 32, nmi), dim3(64), 0, st, dM);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sa_update, dim3(nblk2(mx, 256), nmi), dim3(256), 0, st, dM, ovl);
  } else {
    if (mk_ <= SAK_NMAX16) {
      const size_t lds = (size_t)5 * mk_ * SAK_TJ16 * sizeof(double);
      static bool attr = false;
      if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_sa_solve_kt<SAK_TJ16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(5 * SAK_NMAX16 * SAK_TJ16 * sizeof(double)));
        attr = true;
      }
      hipLaunchKernelGGL(k_sa_solve_kt<SAK_TJ16>, dim3((mj + SAK_TJ16 - 1) / SAK_TJ16, nmi), dim3(32 * SAK_TJ16), lds,
                         st, dM, ovl, mk_);
    } else if (mk_ <= SAK_NMAX) {
      const size_t lds = (size_t)5 * mk_ * SAK_TJ * sizeof(double);
      static bool attr = false;
      if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_sa_solve_kt<SAK_TJ>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(5 * SAK_NMAX * SAK_TJ * sizeof(double)));
        attr = true;
      }
      hipLaunchKernelGGL(k_sa_solve_kt<SAK_TJ>, dim3((mj + SAK_TJ - 1) / SAK_TJ, nmi), dim3(32 * SAK_TJ), lds, st,
                         dM, ovl, mk_);
    } else {
      hipLaunchKernelGGL(k_sa_solve_k, dim3(mj, nmi), dim3(64), 0, st, dM, ovl);
    }
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_SOLVE_K, st);
  hipLaunchKernelGGL(k_turbc_eddy, dim3(nmi), dim3(BC_THREADS), 0, st, dM, ovl);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_EDDY, st);
  return hipSuccess;
}

}  // namespace ot2d
