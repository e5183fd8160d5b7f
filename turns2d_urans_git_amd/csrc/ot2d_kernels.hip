// gfx950 kernels of the OVERTURNS-2D iteration (FP64 plain path).
//
// Kernel map (reference routine -> kernel):
//   time_step        src/time.f:2-107          k_time_step      (all cells)
//   bc               src/bc.f:2-85 (+ bc*)      k_bc             (1 block / instance)
//   rhsup+roeflx+    src/eulerrhs.f:2-446,       k_residual       (32x8 tiles, LDS)
//   muscld_new       821-912; norms step.f:53-115
//   timespectralrhs  src/sources.f:2-27          k_tsrhs
//   rbc              src/bc.f:88-177             k_rbc            (1 block / instance)
//   s*tscale, ilu2d  src/step.f:141-159,         FP64: k_lu_prep, k_lu_sweep<double,RL>
//                    src/lhs.f:470-814            (one wave per instance, skewed rows),
//                                                 k_lu_finish; FP32 operator: k_lu_prep32,
//                                                 k_lu_sweep32 (32-row workgroups),
//                                                 k_lu_finish<float>
//   q update         src/step.f:205-211          fused in k_lu_finish (Euler) or
//   monitor          src/step.f:332-392          k_update_monitor; k_monitor alone
//   initia geometry  src/init.f:2-177, setup.f   k_init_geometry
//   force2d          src/forces.f:111-398        k_forces
//
// Arithmetic follows the reference expressions operand for operand; the
// library is built with -ffp-contract=off so the FP64 path reproduces the
// reference's rounding except where reductions are reordered (norms,
// forces).
#include "ot2d_dev.h"

namespace ot2d {

__device__ __forceinline__ double sq(double a) { return a * a; }
__device__ __forceinline__ double cube(double a) { return a * a * a; }
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

// face speed of a moving grid from the area the face sweeps in one time step
// (src/eulerrhs.f:118-136, 225-243): face between vertices p and m, sgn -1
// for xi faces and +1 for eta faces; second order with the earlier step
// when ntac = 2
__device__ inline double swept_tj(const DevMesh& M, long long p, long long m, double sgn) {
  const double dt = M.c.h;
  double dx1 = M.xv[p] - M.xold[m], dy1 = M.yv[p] - M.yold[m];
  double dx2 = M.xold[p] - M.xv[m], dy2 = M.yold[p] - M.yv[m];
  double tj = sgn * 0.5 * (dx1 * dy2 - dx2 * dy1) / dt;
  if (M.c.ntac == 2 && *M.istep > 1) {
    dx1 = M.xold[p] - M.xole[m]; dy1 = M.yold[p] - M.yole[m];
    dx2 = M.xole[p] - M.xold[m]; dy2 = M.yole[p] - M.yold[m];
    const double temp = sgn * 0.5 * (dx1 * dy2 - dx2 * dy1) / dt;
    tj = 1.5 * tj - 0.5 * temp;
  }
  return tj;
}
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// =================================================================== time_step
// src/time.f:2-107 (one value per cell, all cells incl. halo)
// itn > 0: Newton sub-iteration itn of a time-accurate step, whose dual time
// is dtpseudo(itn) (src/Turns_overset.f:478)
__device__ inline void time_step_cell(const DevMesh& M, long long i, int itn) {
  const DevConst& C = M.c;
  const double oat = time_oat(M);
  const double dualtime = itn > 0 ? C.dtps[itn - 1] : C.dualtime;
  double ibp = (double)imax(M.iblank[i], 0);
  double qj = M.q[M.nq - 1][i];
  if (C.iconstant_cfl == 0) {
    if (C.idual == 0) {
      M.tscale[i] = oat * C.h * ibp * (1.0 + C.jac_factor * (1. - C.timeac) * sqrt(qj)) /
                    (1.0 + (1. - C.timeac) * sqrt(qj));
      M.bt[i] = 1.;
    } else {
      double t = ibp * (1.0 + C.jac_factor * sqrt(qj)) / (1.0 + sqrt(qj));
      if (C.iprecon) t = 100. * ibp * (1.0 + 0.005 * sqrt(qj)) / (1.0 + sqrt(qj));
      t = t * dualtime;
      M.bt[i] = 1. / (1. + t / C.h / oat);
      M.tscale[i] = t / (1. + t / C.h / oat);
    }
    return;
  }
  double q1 = M.q[0][i], q2 = M.q[1][i], q3 = M.q[2][i], q4 = M.q[3][i];
  double u = q2 / q1, v = q3 / q1, e = q4 / q1;
  double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
  double xx = M.xx[i], xy = M.xy[i], yx = M.yx[i], yy = M.yy[i];
  double uu = (u - ug) * xx + (v - vg) * xy;
  double vv = (u - ug) * yx + (v - vg) * yy;
  double dxdx = sq(xx) + sq(xy);
  double dydy = sq(yx) + sq(yy);
  double asq = C.ggm1 * (e - 0.5 * (sq(u) + sq(v)));
  double lam_j = fabs(uu) + sqrt(asq * dxdx);
  double lam_k = fabs(vv) + sqrt(asq * dydy);
  if (C.idual == 0) {
    double t = C.h / (lam_j + lam_k);
    M.tscale[i] = t * ibp;
    M.bt[i] = 1.;
  } else {
    double hnew = dualtime / (lam_j + lam_k);
    double t = hnew * ibp;
    t = t / (1. + t / C.h / oat);
    M.tscale[i] = t;
    M.bt[i] = 1. / (1. + t / C.h / oat);
  }
}

// With ncase > 0 the first workgroup also opens the step of every running
// case (istep0 and the step counter; src/Turns_overset.f:364): the
// boundary conditions launched next are the first readers of istep0.
__global__ __launch_bounds__(256) void k_time_step(const DevMesh* __restrict__ Ms, int* istep0, const int* status,
                                                   int* stepc, int ncase, int itn) {
  if (blockIdx.x == 0 && blockIdx.y == 0)
    for (int c = threadIdx.x; c < ncase; c += 256)
      if (status[c] == 0) {
        istep0[c] += 1;
        stepc[c] += 1;
      }
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  long long n2 = (long long)M.jd * M.kd;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x)
    time_step_cell(M, i, itn);
}

// ====================================================== time-accurate steps
// A time step of a time-accurate run opens here: istep0, the call's step
// counter and the run's time-step counter istep of every running case
// (src/Turns_overset.f:362-365); its Newton sub-iterations follow.
__global__ __launch_bounds__(256) void k_open_step(const int* status, int* istep0, int* stepc, int* istep,
                                                   int ncase) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < ncase && status[c] == 0) {
    istep0[c] += 1;
    stepc[c] += 1;
    istep[c] += 1;
  }
}

// stqol (src/time.f:109-172): the Newton target qnewt (the state, plus the
// BDF2/BDF3 extrapolation once there are earlier levels) and the shift of
// the stored levels, every cell and variable 1..nv
__global__ __launch_bounds__(256) void k_stqol(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const long long n2 = (long long)M.jd * M.kd;
  const int istep = *M.istep, ntac = M.c.ntac, nv = M.nv;
  const bool bdf2 = (ntac == 2 && istep > 1) || (ntac == 3 && istep == 2);
  const bool bdf3 = ntac == 3 && istep > 2;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x)
    for (int n = 0; n < nv; ++n) {
      double* qtn = M.ta + (long long)n * n2;
      double* qtnm1 = M.ta + (long long)(nv + n) * n2;
      double* qnewt = M.ta + (long long)(2 * nv + n) * n2;
      const double q = M.q[n][i], a = qtn[i], b = qtnm1[i];
      double w = q;
      if (bdf2) w = w + (q - a) / 3.;
      if (bdf3) w = w + 7. * (q - a) / 11. - 2. * (a - b) / 11.;
      qnewt[i] = w;
      qtnm1[i] = a;
      qtn[i] = q;
    }
}

// newton (src/time.f:174-209): the time term of the Newton iteration on the
// interior residual, s - (q - qnewt) / (oat h), variables 1..nv
__global__ __launch_bounds__(256) void k_newton(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int nj = M.je - M.js + 1, nk = M.ke - M.ks + 1;
  const long long ncell = (long long)nj * nk;
  const long long n2 = (long long)M.jd * M.kd;
  const double tac = 1. / (time_oat(M) * M.c.h);
  for (long long w = blockIdx.x * 256ll + threadIdx.x; w < ncell; w += 256ll * gridDim.x) {
    const long long i = cell(M, M.js + (int)(w % nj), M.ks + (int)(w / nj));
    for (int n = 0; n < M.nv; ++n) {
      const double* qnewt = M.ta + (long long)(2 * M.nv + n) * n2;
      M.s[n][i] = M.s[n][i] - (M.q[n][i] - qnewt[i]) * tac;
    }
  }
}

// n doubles from src to dst (a graph holds no copy nodes)
__global__ __launch_bounds__(256) void k_copy_doubles(double* __restrict__ dst, const double* __restrict__ src,
                                                      int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ======================================================================= BCs
struct QV {
  double *q1, *q2, *q3, *q4, *qj;
};
__device__ inline QV qview(const DevMesh& M) {
  QV v;
  v.q1 = M.q[0]; v.q2 = M.q[1]; v.q3 = M.q[2]; v.q4 = M.q[3]; v.qj = M.q[M.nq - 1];
  return v;
}

// pressure (times J) of a cell as the reference extrapolates it
__device__ inline double p_of(const QV& v, double gm1, long long i) {
  return gm1 * (v.q4[i] - 0.5 * (sq(v.q2[i]) + sq(v.q3[i])) / v.q1[i]) * v.qj[i];
}

// bcwall + bctany + uv for one j of a k=const wall (src/bc.f:1669-1801,
// 1328-1415, 2224-2264), including the 30-step ramp of the same cell.
__device__ inline void bcwall_cell(const DevMesh& M, const QV& v, int j, int k, int k1, int k2,
                                   int invsc, double scal, double rscal) {
  const DevConst& C = M.c;
  long long i = cell(M, j, k), i1 = cell(M, j, k1), i2 = cell(M, j, k2);
  double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
  double ug1 = M.has_gv ? M.ug[i1] : 0.0, vg1 = M.has_gv ? M.vg[i1] : 0.0;
  double xx = M.xx[i], xy = M.xy[i], yx = M.yx[i], yy = M.yy[i];
  double xx1 = M.xx[i1], xy1 = M.xy[i1], yx1 = M.yx[i1], yy1 = M.yy[i1];
  double u, vv;
  if (invsc) {
    double f = 1.0;
    double u1 = (v.q2[i1] * xx + v.q3[i1] * xy) / v.q1[i1];
    double u2 = (v.q2[i2] * xx + v.q3[i2] * xy) / v.q1[i2];
    u = (1. + f) * u1 - f * u2;
    double v1 = (v.q2[i1] * yx + v.q3[i1] * yy) / v.q1[i1];
    vv = ug * yx + vg * yy + ug1 * yx1 + vg1 * yy1 - v1;
  } else {
    double u1 = (v.q2[i1] * xx + v.q3[i1] * xy) / v.q1[i1];
    double v1 = (v.q2[i1] * yx + v.q3[i1] * yy) / v.q1[i1];
    u = ug * xx + vg * xy + ug1 * xx1 + vg1 * xy1 - u1;
    vv = ug * yx + vg * yy + ug1 * yx1 + vg1 * yy1 - v1;
  }
  double t11 = yy, t12 = -xy, t21 = -yx, t22 = xx;
  double vol = 1. / (t22 * t11 - t12 * t21);
  double uwall = (t11 * u + t12 * vv) * vol;
  double vwall = (t21 * u + t22 * vv) * vol;
  double foso = invsc ? 1.0 : 0.0;
  double rho1 = v.q1[i1] * v.qj[i1];
  double rho2 = v.q1[i2] * v.qj[i2];
  double p1 = p_of(v, C.gm1, i1);
  double p2 = p_of(v, C.gm1, i2);
  double qj = v.qj[i];
  double r = ((1. + foso) * rho1 - foso * rho2) / qj;
  double m2 = uwall * r;
  double m3 = vwall * r;
  double press = (1. + foso) * p1 - foso * p2;
  double e = press / (C.gm1 * qj) + 0.5 * (sq(m2) + sq(m3)) / r;
  double ajacinv = 1. / qj;
  v.q1[i] = r * scal + rscal * C.rinf * ajacinv;
  v.q2[i] = m2 * scal + rscal * C.rinf * C.uinf * ajacinv;
  v.q3[i] = m3 * scal + rscal * C.rinf * C.vinf * ajacinv;
  v.q4[i] = e * scal + rscal * C.einf * ajacinv;
}

// bcout_trkl first layer (src/bc.f:694-983)
__device__ inline void trkl_cell(const DevMesh& M, const QV& v, long long i, long long i1,
                                 long long i2, double xxn, double xyn) {
  const DevConst& C = M.c;
  double gm1 = C.gm1, gamma = C.gamma, Mp = C.Mp;
  double foso = 1.0;
  double qext[4], qbar[4], delq[4], delw[4];
  double* qs[4] = {v.q1, v.q2, v.q3, v.q4};
#pragma unroll
  for (int n = 0; n < 4; ++n) qext[n] = (1. + foso) * qs[n][i1] * v.qj[i1] - foso * qs[n][i2] * v.qj[i2];
  qbar[0] = 0.5 * (qext[0] + C.rinf);
  qbar[1] = 0.5 * (qext[1] + C.rinf * C.uinf);
  qbar[2] = 0.5 * (qext[2] + C.rinf * C.vinf);
  qbar[3] = 0.5 * (qext[3] + C.einf);
  delq[0] = 0.5 * (C.rinf - qext[0]);
  delq[1] = 0.5 * (C.rinf * C.uinf - qext[1]);
  delq[2] = 0.5 * (C.rinf * C.vinf - qext[2]);
  delq[3] = 0.5 * (C.einf - qext[3]);
  double rho = qbar[0];
  double rrho = 1. / rho;
  double uu = qbar[1] * rrho, vv = qbar[2] * rrho, e = qbar[3] * rrho;
  double uv2 = 0.5 * (uu * uu + vv * vv);
  double cjkl = sqrt(C.ggm1 * (e - uv2));
  double c2i = 1. / (cjkl * cjkl);
  double ge = gamma * e - gm1 * uv2;
  double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
  double qq = (uu - ug) * xxn + (vv - vg) * xyn;
  double bt = M.bt[i];
  double bSq = sq(Mp) / (bt - sq(Mp) * (bt - 1));
  double Z = sqrt((1. - bSq) * qq * (1. - bSq) * qq + 4. * bSq * cjkl * cjkl);
  double Zi = 1. / Z;
  double R = 0.5 * ((1. - bSq) * qq + Z);
  double S = 0.5 * ((1. - bSq) * qq - Z);
  double sq0 = fsign(1., qq);
  double sR = fsign(1., R + bSq * qq), sR2 = fsign(1., R + qq * bSq);
  double sS = fsign(1., S + bSq * qq), sS2 = fsign(1., S + qq * bSq);
  delw[0] = (1. - uv2 * gm1 * c2i) * sq0 * delq[0];
  delw[0] = delw[0] + uu * c2i * gm1 * sq0 * delq[1];
  delw[0] = delw[0] + vv * c2i * gm1 * sq0 * delq[2];
  delw[0] = delw[0] - gm1 * c2i * sq0 * delq[3];
  delw[1] = (vv * xxn - uu * xyn) * sq0 * delq[0];
  delw[1] = delw[1] + xyn * sq0 * delq[1];
  delw[1] = delw[1] - xxn * sq0 * delq[2];
  delw[2] = -(Zi * (uv2 * gm1 * c2i * S + bSq * qq) * sR * delq[0]);
  delw[2] = delw[2] + (xxn * bSq + S * uu * gm1 * c2i) * Zi * sR * delq[1];
  delw[2] = delw[2] + (xyn * bSq + S * vv * gm1 * c2i) * Zi * sR2 * delq[2];
  delw[2] = delw[2] - S * gm1 * c2i * Zi * sR2 * delq[3];
  delw[3] = Zi * (uv2 * gm1 * c2i * R + qq * bSq) * sS * delq[0];
  delw[3] = delw[3] - (xxn * bSq + R * uu * gm1 * c2i) * Zi * sS * delq[1];
  delw[3] = delw[3] - (xyn * bSq + R * vv * gm1 * c2i) * Zi * sS2 * delq[2];
  delw[3] = delw[3] + R * gm1 * c2i * Zi * sS2 * delq[3];
  delq[0] = delw[0] + delw[2] + delw[3];
  delq[1] = uu * delw[0] + xyn * delw[1] + (uu + xxn * R / bSq) * delw[2];
  delq[1] = delq[1] + (uu + xxn * S / bSq) * delw[3];
  delq[2] = vv * delw[0] - xxn * delw[1] + (vv + xyn * R / bSq) * delw[2];
  delq[2] = delq[2] + (vv + xyn * S / bSq) * delw[3];
  delq[3] = uv2 * delw[0] + (uu * xyn - vv * xxn) * delw[1];
  delq[3] = delq[3] + (ge + R * qq / bSq) * delw[2];
  delq[3] = delq[3] + (ge + S * qq / bSq) * delw[3];
  double qj = v.qj[i];
#pragma unroll
  for (int n = 0; n < 4; ++n) qs[n][i] = (qbar[n] - delq[n]) / qj;
}

// periodic source index (src/bc.f:192-262)
__device__ inline int periodic_src(int idir, int js, int je, int ks, int ke, int jd, int kd, int jk) {
  if (idir == 1) { int jj = je - js + 1; return jd - 2 * jj + (jk - js + 1); }
  if (idir == -1) { int jj = je - js + 1; return 1 + 2 * jj - (je - jk + 1); }
  if (idir == 2) { int kk = ke - ks + 1; return kd - 2 * kk + (jk - ks + 1); }
  int kk = ke - ks + 1;
  return 1 + 2 * kk - (ke - jk + 1);
}

// One boundary segment, cells distributed over the block.  Every segment
// type only reads cells outside the rows/columns it writes, so the cells of
// one segment are independent; segments are separated by barriers to keep
// the reference's sequential order (src/bc.f:19-82).
__device__ void bc_segment(const DevMesh& M, const DevBC& b, int istep0) {
  const DevConst& C = M.c;
  QV v = qview(M);
  int js = b.js, je = b.je, ks = b.ks, ke = b.ke, idir = b.idir;
  int iadd = idir >= 0 ? 1 : -1, iadir = idir >= 0 ? idir : -idir;
  int t = threadIdx.x, nt = blockDim.x;
  switch (b.type) {
    case 3: case 4: case 5: {
      if (iadir != 2) return;
      int invsc = b.type == 3 ? 1 : (b.type == 4 ? 0 : C.invisc);
      int k = (idir == 2) ? ke : ks;
      double tt = (double)istep0 / 30.;
      if (tt > 1. || C.iread > 0) tt = 1.;
      double scal = (10. - 15. * tt + 6. * tt * tt) * cube(tt);
      double rscal = 1. - scal;
      // one halo layer: ks == ke, the ramp touches exactly the wall row
      for (int j = js + t; j <= je; j += nt) bcwall_cell(M, v, j, k, k + iadd, k + 2 * iadd, invsc, scal, rscal);
      return;
    }
    case 10: {
      double foso = 0.0;
      double* qs[4] = {v.q1, v.q2, v.q3, v.q4};
      if (iadir == 1) {
        int j = (idir == 1) ? je : js;  // one layer
        int j1 = j + iadd, j2 = j1 + iadd;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j1, k), i2 = cell(M, j2, k);
          double s1 = v.qj[i1] / v.qj[i], s2 = v.qj[i2] / v.qj[i];
          for (int n = 0; n < 4; ++n) qs[n][i] = (1. + foso) * qs[n][i1] * s1 - foso * qs[n][i2] * s2;
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks;
        int k1 = k + iadd, k2 = k1 + iadd;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j, k1), i2 = cell(M, j, k2);
          double s1 = v.qj[i1] / v.qj[i], s2 = v.qj[i2] / v.qj[i];
          for (int n = 0; n < 4; ++n) qs[n][i] = (1. + foso) * qs[n][i1] * s1 - foso * qs[n][i2] * s2;
        }
      }
      return;
    }
    case 11: {
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = (idir == 1) ? je + 1 : js - 1;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j1, k);
          double sc = v.qj[i1] / v.qj[i];
          v.q1[i] = v.q1[i1] * sc; v.q2[i] = -v.q2[i1] * sc;
          v.q3[i] = v.q3[i1] * sc; v.q4[i] = v.q4[i1] * sc;
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j, k1);
          double sc = v.qj[i1] / v.qj[i];
          v.q1[i] = v.q1[i1] * sc; v.q2[i] = v.q2[i1] * sc;
          v.q3[i] = -v.q3[i1] * sc; v.q4[i] = v.q4[i1] * sc;
        }
      }
      return;
    }
    case 22: {
      if (iadir == 1) {
        int ncell = (je - js + 1) * (ke - ks + 1);
        for (int w = t; w < ncell; w += nt) {
          int j = js + w % (je - js + 1), k = ks + w / (je - js + 1);
          int jc = periodic_src(idir, js, je, ks, ke, M.jd, M.kd, j);
          long long i = cell(M, j, k), i1 = cell(M, jc, k);
          double sc = v.qj[i1] / v.qj[i];
          v.q1[i] = v.q1[i1] * sc; v.q2[i] = v.q2[i1] * sc; v.q3[i] = v.q3[i1] * sc; v.q4[i] = v.q4[i1] * sc;
        }
      } else {
        int ncell = (je - js + 1) * (ke - ks + 1);
        for (int w = t; w < ncell; w += nt) {
          int j = js + w % (je - js + 1), k = ks + w / (je - js + 1);
          int kc = periodic_src(idir, js, je, ks, ke, M.jd, M.kd, k);
          long long i = cell(M, j, k), i1 = cell(M, j, kc);
          double sc = v.qj[i1] / v.qj[i];
          v.q1[i] = v.q1[i1] * sc; v.q2[i] = v.q2[i1] * sc; v.q3[i] = v.q3[i1] * sc; v.q4[i] = v.q4[i1] * sc;
        }
      }
      return;
    }
    case 51: {
      if (iadir != 2) return;
      int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
      for (int j = js + t; j <= je; j += nt) {
        int jj = M.jd - j + 1;
        long long a = cell(M, j, k), bb = cell(M, jj, k1), cc = cell(M, jj, k), d = cell(M, j, k1);
        double scale1 = v.qj[bb] / v.qj[a];
        double scale2 = v.qj[d] / v.qj[cc];
        double a1 = v.q1[bb] * scale1, a2 = v.q2[bb] * scale1, a3 = v.q3[bb] * scale1, a4 = v.q4[bb] * scale1;
        double c1 = v.q1[d] * scale2, c2 = v.q2[d] * scale2, c3 = v.q3[d] * scale2, c4 = v.q4[d] * scale2;
        v.q1[a] = a1; v.q2[a] = a2; v.q3[a] = a3; v.q4[a] = a4;
        v.q1[cc] = c1; v.q2[cc] = c2; v.q3[cc] = c3; v.q4[cc] = c4;
      }
      return;
    }
    case 46: {
      int ncell = (je - js + 1) * (ke - ks + 1);
      for (int w = t; w < ncell; w += nt) {
        int j = js + w % (je - js + 1), k = ks + w / (je - js + 1);
        long long i = cell(M, j, k);
        double rj = 1. / v.qj[i];
        double r = C.rinf * rj, m2 = C.rinf * C.uinf * rj, m3 = C.rinf * C.vinf * rj;
        v.q1[i] = r; v.q2[i] = m2; v.q3[i] = m3;
        v.q4[i] = C.pinf / C.gm1 * rj + 0.5 * (sq(m2) + sq(m3)) / r;
      }
      return;
    }
    case 47: {
      if (iadir == 1) {
        int j = (idir == 1) ? je : js;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k);
          double snorm = -1. * iadd / sqrt(sq(M.xx[i]) + sq(M.xy[i]));
          trkl_cell(M, v, i, cell(M, j + iadd, k), cell(M, j + 2 * iadd, k), M.xx[i] * snorm, M.xy[i] * snorm);
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k);
          double snorm = -1. * iadd / sqrt(sq(M.yx[i]) + sq(M.yy[i]));
          trkl_cell(M, v, i, cell(M, j, k + iadd), cell(M, j, k + 2 * iadd), M.yx[i] * snorm, M.yy[i] * snorm);
        }
      }
      return;
    }
    default:
      return;
  }
}

// zero s on the halo ring (the reference clears all of s at step start,
// src/step.f:41; interior s is fully rewritten by k_residual)
__device__ void zero_halo_s(const DevMesh& M) {
  int jd = M.jd, kd = M.kd;
  int nring = 2 * jd + 2 * (kd - 2);
  for (int w = threadIdx.x; w < nring; w += blockDim.x) {
    int j, k;
    if (w < jd) { j = w + 1; k = 1; }
    else if (w < 2 * jd) { j = w - jd + 1; k = kd; }
    else if (w < 2 * jd + kd - 2) { j = 1; k = w - 2 * jd + 2; }
    else { j = jd; k = w - 2 * jd - (kd - 2) + 2; }
    long long i = cell(M, j, k);
    for (int n = 0; n < M.nv; ++n) M.s[n][i] = 0.0;
  }
}

// bc (src/bc.f:2-85): the segments of one phase of every instance, one
// workgroup per segment (segments of a phase touch disjoint cells; phases
// run in the reference's segment order); the first workgroup of phase 0
// also clears s on the halo ring
__global__ __launch_bounds__(BC_THREADS) void k_bc(const DevMesh* __restrict__ Ms, int zero_s, int phase) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int ib = blockIdx.x;
  if (zero_s && phase == 0 && ib == 0) zero_halo_s(M);
  if (ib >= M.nbc || M.bc[ib].phase != phase) return;
  bc_segment(M, M.bc[ib], *M.istep0);
}

// =================================================================== residual
// MUSCL kappa=1/3 with Koren's differentiable limiter for cell i of a line
// (src/eulerrhs.f:856-872): returns ql(i), qr(i).
__device__ __forceinline__ void muscl_cell(double fm, double f0, double fp, int ibm, int ib0, int ibp,
                                           double th, double qt, double eps, double& ql, double& qr) {
  double thm = 1.0 - th, thp = 1.0 + th;
  double f2i = fp - f0, f2i1 = f0 - fm;
  double a1 = 3.0 * (f2i * f2i1 + eps);
  double a2 = 2.0 * sq(f2i - f2i1) + a1;
  double f3i = a1 / a2;
  double f3qt = qt * f3i * (double)ib0 * (double)ibp * (double)ibm;
  ql = f0 + f3qt * (thm * f2i1 + thp * f2i);
  qr = f0 - f3qt * (thp * f2i1 + thm * f2i);
}

// Roe flux with the smooth entropy fix (src/eulerrhs.f:283-446) for one face
// with left state l, right state r (rho,u,v,p), face vector (xa,ya), tj.
__device__ __forceinline__ void roe_face(const DevConst& C, const double* l, const double* r, double xa,
                                         double ya, double tj, double* f) {
  double gm1 = C.gm1;
  double eps = 1.e-6;
  double rlft = l[0], ulft = l[1], vlft = l[2], plft = l[3];
  double rlfti = 1.0 / rlft;
  double rulft = rlft * ulft, rvlft = rlft * vlft;
  double uvl = 0.5 * (ulft * ulft + vlft * vlft);
  double elft = plft / gm1 + rlft * uvl;
  double hlft = (elft + plft) * rlfti;
  double clft = sqrt(gm1 * (hlft - uvl));
  double rrht = r[0], urht = r[1], vrht = r[2], prht = r[3];
  double rrhti = 1.0 / rrht;
  double rurht = rrht * urht, rvrht = rrht * vrht;
  double uvr = 0.5 * (urht * urht + vrht * vrht);
  double erht = prht / gm1 + rrht * uvr;
  double hrht = (erht + prht) * rrhti;
  double crht = sqrt(gm1 * (hrht - uvr));
  double rat = sqrt(rrht * rlfti);
  double rati = 1.0 / (rat + 1.);
  double rav = rat * rlft;
  double uav = (rat * urht + ulft) * rati;
  double vav = (rat * vrht + vlft) * rati;
  double hav = (rat * hrht + hlft) * rati;
  double uv = 0.5 * (uav * uav + vav * vav);
  double cav = sqrt(gm1 * (hav - uv));
  double aq1 = rrht - rlft, aq2 = urht - ulft, aq3 = vrht - vlft, aq4 = prht - plft;
  double ri1 = xa, ri2 = ya, ri3 = tj;
  double rr2 = ri1 * ri1 + ri2 * ri2;
  double rr = sqrt(rr2);
  double r0 = 1.0 / rr;
  double r1 = ri1 * r0, r2 = ri2 * r0, r3 = ri3 * r0;
  double uu = r1 * uav + r2 * vav + r3;
  double c2 = cav * cav;
  double c2i = 1.0 / c2;
  double auu = fabs(uu), aupc = fabs(uu + cav), aumc = fabs(uu - cav);
  double uulft = r1 * ulft + r2 * vlft + r3;
  double uurht = r1 * urht + r2 * vrht + r3;
  double upclft = uulft + clft, upcrht = uurht + crht;
  double umclft = uulft - clft, umcrht = uurht - crht;
  double dauu = 4. * (uurht - uulft) + eps;
  double dauus = fmax(dauu, 0.0);
  if (auu <= 0.5 * dauus) auu = sq(auu) / dauu + 0.25 * dauu;
  double daupc = 4. * (upcrht - upclft) + eps;
  double daupcs = fmax(daupc, 0.0);
  if (aupc <= 0.5 * daupcs) aupc = sq(aupc) / daupc + 0.25 * daupc;
  double daumc = 4. * (umcrht - umclft) + eps;
  double daumcs = fmax(daumc, 0.0);
  if (aumc <= 0.5 * daumcs) aumc = sq(aumc) / daumc + 0.25 * daumc;
  double rcav = rav * cav;
  double aquu = uurht - uulft;
  double c2ih = 0.5 * c2i;
  double ruuav = auu * rav;
  double b1 = auu * (aq1 - c2i * aq4);
  double b2 = c2ih * aupc * (aq4 + rcav * aquu);
  double b3 = c2ih * aumc * (aq4 - rcav * aquu);
  double b4 = b1 + b2 + b3;
  double b5 = cav * (b2 - b3);
  double b6 = ruuav * (aq2 - r1 * aquu);
  double b7 = ruuav * (aq3 - r2 * aquu);
  aq1 = b4;
  aq2 = uav * b4 + r1 * b5 + b6;
  aq3 = vav * b4 + r2 * b5 + b7;
  aq4 = hav * b4 + (uu - r3) * b5 + uav * b6 + vav * b7 - c2 * b1 / gm1;
  double aj = 0.5 * rr;
  double plar = plft + prht;
  double eplft = elft + plft, eprht = erht + prht;
  double fssub = 0.0;
  f[0] = aj * (rlft * uulft + rrht * uurht - aq1) - fssub * C.rinf;
  f[1] = aj * (rulft * uulft + rurht * uurht + r1 * plar - aq2) - fssub * C.rinf * C.uinf;
  f[2] = aj * (rvlft * uulft + rvrht * uurht + r2 * plar - aq3) - fssub * C.rinf * C.vinf;
  f[3] = aj * (eplft * uulft + eprht * uurht - r3 * plar - aq4) - fssub * C.einf;
}

// Reconstructed left/right states at the face between line cells i and i+1
// of a line with n cells (1-based) held in LDS (stride st between cells).
// First order at the line ends exactly as muscld_new (is=1, imax=n).
__device__ __forceinline__ void face_states(const double* sf0, int stv, int stc, const int* sib, int stci,
                                            int c0, int i, int n, double th, double qt, double eps,
                                            double* l, double* r) {
  // c0 = LDS column of line cell i
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double* f = sf0 + v * stv;
    double fl0 = f[c0 * stc], fr0 = f[(c0 + 1) * stc];
    if (qt == 0.0) { l[v] = fl0; r[v] = fr0; continue; }
    double dummy;
    if (i == 1) l[v] = fl0;
    else muscl_cell(f[(c0 - 1) * stc], fl0, fr0, sib[(c0 - 1) * stci], sib[c0 * stci], sib[(c0 + 1) * stci],
                    th, qt, eps, l[v], dummy);
    if (i + 1 == n) r[v] = fr0;
    else muscl_cell(fl0, fr0, f[(c0 + 2) * stc], sib[c0 * stci], sib[(c0 + 1) * stci], sib[(c0 + 2) * stci],
                    th, qt, eps, dummy, r[v]);
  }
}

// ---- WENO5 (src/eulerrhs.f:1011-1230) as rhsup calls it: ibmin = ibmax = 1,
// so the fmin/fmax branches are dead.  For a line of n cells the reference
// leaves ql(1) and qr(n) from its end formulas, ql and qr of cells 2 and n-1
// and ql of cell 3 from SONIC-A slopes, and WENO5 (blended to first order
// next to blanked cells) elsewhere; the functions below evaluate each case
// from the cells it reads, with the reference's operation order.
__device__ __forceinline__ double ammd(double fl, double fr) {
  double a = fabs(fl), b = fabs(fr);
  return 0.5 * (copysign(1.0, fl) + copysign(1.0, fr)) * (a < b ? a : b);
}

__device__ __forceinline__ double weno5(double a, double b, double c, double d, double e) {
  const double b1 = 13. / 12., b2 = 1. / 6., epsw = 1.e-6;
  double djm1 = a - 2. * b + c, ejm1 = a - 4. * b + 3. * c;
  double dj = b - 2. * c + d, ej = b - d;
  double djp1 = c - 2. * d + e, ejp1 = 3. * c - 4. * d + e;
  double dis0 = b1 * djm1 * djm1 + 0.25 * ejm1 * ejm1 + epsw;
  double dis1 = b1 * dj * dj + 0.25 * ej * ej + epsw;
  double dis2 = b1 * djp1 * djp1 + 0.25 * ejp1 * ejp1 + epsw;
  double q30 = 2. * a - 7. * b + 11. * c;
  double q31 = -b + 5. * c + 2. * d;
  double q32 = 2. * c + 5. * d - e;
  double a1ba0 = 6. * (dis0 / dis1), a2ba0 = 3. * (dis0 / dis2);
  double w0 = 1. / (1. + a1ba0 + a2ba0);
  double w1 = a1ba0 * w0;
  double w2 = 1. - w0 - w1;
  return b2 * (w0 * q30 + w1 * q31 + w2 * q32);
}

// SONIC-A slope from f1(i), f1(i-1) and the limited curvatures f2m(i), f2m(i-1)
__device__ __forceinline__ double sonic(double f1i, double f1m, double f2mi, double f2mm) {
  double at = f1i - 0.5 * f2mi, ati1 = f1m + 0.5 * f2mm;
  double s1 = ammd(f1i, f1m), t1 = ammd(at, ati1);
  double a = 0.5 * fabs(at + ati1), b = 2. * fabs(s1), c = fabs(t1);
  double mx = b > c ? b : c;
  return copysign(1.0, t1) * (a < mx ? a : mx);
}

// slope at cell 2 from f0(1..4) (f2(1) extrapolated)
__device__ __forceinline__ double sonic_lo(double x1, double x2, double x3, double x4, double& f1_1, double& f2m_1) {
  double f11 = x2 - x1, f12 = x3 - x2, f13 = x4 - x3;
  double f22 = f12 - f11, f23 = f13 - f12;
  double f21 = 2. * f22 - f23;
  f1_1 = f11;
  f2m_1 = ammd(f21, f22);
  return sonic(f12, f11, ammd(f22, f23), f2m_1);
}

// slope at cell n-1 from f0(n-3..n) (f2(n) extrapolated)
__device__ __forceinline__ double sonic_hi(double y0, double y1, double y2, double y3, double& f1_n1, double& f2m_n1) {
  double fa = y1 - y0, fb = y2 - y1, fc = y3 - y2;    // f1(n-3), f1(n-2), f1(n-1)
  double f2a = fb - fa, f2b = fc - fb;                // f2(n-2), f2(n-1)
  double f2n = 2. * f2b - f2a;
  f1_n1 = fc;
  f2m_n1 = ammd(f2b, f2n);
  return sonic(fc, fb, f2m_n1, ammd(f2a, f2b));
}

// slope at cell i from f0(i-2..i+2) (no extrapolated curvature)
__device__ __forceinline__ double sonic_mid(double z0, double z1, double z2, double z3, double z4) {
  double fa = z1 - z0, fb = z2 - z1, fc = z3 - z2, fd = z4 - z3;   // f1(i-2..i+1)
  double f2a = fb - fa, f2b = fc - fb, f2c = fd - fc;             // f2(i-1..i+1)
  return sonic(fc, fb, ammd(f2b, f2c), ammd(f2a, f2b));
}

// left state of cell i and right state of cell i+1 (the two sides of face
// i) of a line of n cells, from x[0..5] = f0(i-2..i+3), ib[0..5] = |iblank|
__device__ __forceinline__ void weno_face(const double* x, const int* ib, int i, int n, double& ql, double& qr) {
  double f1, f2m;
  if (i == 1) {
    sonic_lo(x[2], x[3], x[4], x[5], f1, f2m);
    double at = f1 - 0.5 * f2m;
    ql = x[2] + 0.5 * ammd(at, 2. * f1);
  } else if (i == 2) {
    ql = x[2] + 0.5 * sonic_lo(x[1], x[2], x[3], x[4], f1, f2m);
  } else if (i == 3 || i == n - 1) {
    ql = x[2] + 0.5 * (i == 3 ? sonic_mid(x[0], x[1], x[2], x[3], x[4]) : sonic_hi(x[0], x[1], x[2], x[3], f1, f2m));
  } else {
    double w = weno5(x[0], x[1], x[2], x[3], x[4]);
    double bv = (double)(ib[2] * ib[1] * ib[3] * ib[0] * ib[4]);
    ql = w * bv + (1. - bv) * x[2];
  }
  const int m = i + 1;
  if (m == n) {
    sonic_hi(x[0], x[1], x[2], x[3], f1, f2m);
    double at1 = f1 + 0.5 * f2m;
    qr = x[3] - 0.5 * ammd(at1, 2. * f1);
  } else if (m == 2) {
    qr = x[3] - 0.5 * sonic_lo(x[2], x[3], x[4], x[5], f1, f2m);
  } else if (m == n - 1) {
    qr = x[3] - 0.5 * sonic_hi(x[1], x[2], x[3], x[4], f1, f2m);
  } else {
    double w = weno5(x[5], x[4], x[3], x[2], x[1]);
    double bv = (double)(ib[3] * ib[2] * ib[4] * ib[1] * ib[5]);
    qr = w * bv + (1. - bv) * x[3];
  }
}

// face_states for ilim = 4: c0 = LDS column of line cell i; the tile holds
// three halo cells on each side
__device__ __forceinline__ void face_states_weno(const double* sf0, int stv, int stc, const int* sib, int stci,
                                                 int c0, int i, int n, double qt, double* l, double* r) {
  int ib[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) ib[d] = sib[(c0 - 2 + d) * stci];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double* f = sf0 + v * stv;
    if (qt == 0.0) { l[v] = f[c0 * stc]; r[v] = f[(c0 + 1) * stc]; continue; }
    double x[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) x[d] = f[(c0 - 2 + d) * stc];
    weno_face(x, ib, i, n, l[v], r[v]);
  }
}


// Residual of one 32x8 tile: rhsup for nei=1 (src/eulerrhs.f:2-280), fluxes
// of the tile's faces computed once into LDS, plus the residual-norm partials
// of step (src/step.f:60-93).  Writes s(:,:,1:4) of the tile's interior cells.
// H: halo cells the tile loads on each side (2 for MUSCL, 3 for WENO5)
template <int H>
__global__ __launch_bounds__(RES_THREADS) void k_residual(const DevMesh* __restrict__ Ms,
                                                          const ResTile* __restrict__ tiles, int do_norms) {
  constexpr int TJ = RES_TJ, TK = RES_TK, LJ = TJ + 2 * H, LK = TK + 2 * H;
  __shared__ double sf[4][LK][LJ];   // primitives rho, u, v, p
  __shared__ int sib[LK][LJ];
  __shared__ double sxv[TK + 1][TJ + 1], syv[TK + 1][TJ + 1];
  __shared__ double sF[4][TK][TJ + 1];  // j-face fluxes
  __shared__ double sG[4][TK + 1][TJ];  // k-face fluxes
  __shared__ double red[4][RES_THREADS];
  const ResTile T = tiles[blockIdx.x];
  const DevMesh& M = Ms[T.mesh];
  const DevConst& C = M.c;
  const int tid = threadIdx.x;
  const int jd = M.jd, kd = M.kd;
  const int j0 = T.j0, k0 = T.k0;
  QV v = qview(M);
  // ---- load primitives of the tile + H-cell halo (corners unused)
  for (int w = tid; w < LK * LJ; w += RES_THREADS) {
    int r = w / LJ, cc = w % LJ;
    int j = j0 - H + cc, k = k0 - H + r;
    bool row_in = (r >= H && r < H + TK), col_in = (cc >= H && cc < H + TJ);
    if (!row_in && !col_in) continue;
    int jj = j < 1 ? 1 : (j > jd ? jd : j);
    int kk = k < 1 ? 1 : (k > kd ? kd : k);
    long long i = cell(M, jj, kk);
    double q1 = v.q1[i], q2 = v.q2[i], q3 = v.q3[i], q4 = v.q4[i], qj = v.qj[i];
    double rhoi = 1.0 / q1;
    double ppp = C.gm1 * (q4 - 0.5 * (sq(q2) + sq(q3)) / q1);
    sf[0][r][cc] = q1 * qj;
    sf[1][r][cc] = q2 * rhoi;
    sf[2][r][cc] = q3 * rhoi;
    sf[3][r][cc] = ppp * qj;
    int ib = M.iblank[i];
    sib[r][cc] = ib < 0 ? -ib : ib;
  }
  for (int w = tid; w < (TK + 1) * (TJ + 1); w += RES_THREADS) {
    int r = w / (TJ + 1), cc = w % (TJ + 1);
    int jv = j0 - 1 + cc, kv = k0 - 1 + r;
    jv = jv > M.jmax ? M.jmax : jv;
    kv = kv > M.kmax ? M.kmax : kv;
    long long iv = vert(M, jv, kv);
    sxv[r][cc] = M.xv[iv];
    syv[r][cc] = M.yv[iv];
  }
  __syncthreads();
  const double th = C.th, qt = C.qt;
  // ---- xi (j-direction) faces: face fi between cells fi and fi+1, rows k0..k0+TK-1
  for (int w = tid; w < TK * (TJ + 1); w += RES_THREADS) {
    int r = w / (TJ + 1), cc = w % (TJ + 1);
    int fi = j0 - 1 + cc, k = k0 + r;
    if (fi > M.je || k > M.ke) continue;
    double l[4], rr[4], f[4];
    // LDS column of cell fi: fi - (j0 - H)
    if (H == 3 && C.ilim == 4)
      face_states_weno(&sf[0][r + H][0], LK * LJ, 1, &sib[r + H][0], 1, fi - (j0 - H), fi, jd, qt, l, rr);
    else
      face_states(&sf[0][r + H][0], LK * LJ, 1, &sib[r + H][0], 1, fi - (j0 - H), fi, jd, th, qt, M.epsj, l, rr);
    // face vector at vertex column jv = fi, rows kv-1..kv with kv = k
    double xa = syv[r + 1][cc] - syv[r][cc];
    double ya = -(sxv[r + 1][cc] - sxv[r][cc]);
    double tj = 0.0;
    if (C.iunst > 0) {
      tj = swept_tj(M, vert(M, fi, k), vert(M, fi, k - 1), -1.);
    } else if (M.has_gv) {
      long long a = vert(M, fi, k), b = vert(M, fi, k - 1);
      tj = -0.5 * (M.ugv[a] + M.ugv[b]) * xa - 0.5 * (M.vgv[a] + M.vgv[b]) * ya;
    }
    roe_face(C, l, rr, xa, ya, tj, f);
#pragma unroll
    for (int n = 0; n < 4; ++n) sF[n][r][cc] = f[n];
  }
  // ---- eta (k-direction) faces: face fk between cells fk and fk+1, columns j0..j0+TJ-1
  for (int w = tid; w < (TK + 1) * TJ; w += RES_THREADS) {
    int r = w / TJ, cc = w % TJ;
    int fk = k0 - 1 + r, j = j0 + cc;
    if (fk > M.ke || j > M.je) continue;
    double l[4], rr[4], f[4];
    if (H == 3 && C.ilim == 4)
      face_states_weno(&sf[0][0][cc + H], LK * LJ, LJ, &sib[0][cc + H], LJ, fk - (k0 - H), fk, kd, qt, l, rr);
    else
      face_states(&sf[0][0][cc + H], LK * LJ, LJ, &sib[0][cc + H], LJ, fk - (k0 - H), fk, kd, th, qt, M.epsk, l, rr);
    // vertex row kv = fk, columns jv-1..jv with jv = j
    double xa = -(syv[r][cc + 1] - syv[r][cc]);
    double ya = (sxv[r][cc + 1] - sxv[r][cc]);
    double tj = 0.0;
    if (C.iunst > 0) {
      tj = swept_tj(M, vert(M, j, fk), vert(M, j - 1, fk), 1.);
    } else if (M.has_gv) {
      long long a = vert(M, j, fk), b = vert(M, j - 1, fk);
      tj = -0.5 * (M.ugv[a] + M.ugv[b]) * xa - 0.5 * (M.vgv[a] + M.vgv[b]) * ya;
    }
    roe_face(C, l, rr, xa, ya, tj, f);
#pragma unroll
    for (int n = 0; n < 4; ++n) sG[n][r][cc] = f[n];
  }
  __syncthreads();
  // ---- flux balance
  int cx = tid % TJ, ry = tid / TJ;
  int j = j0 + cx, k = k0 + ry;
  double sv[4] = {0.0, 0.0, 0.0, 0.0};
  bool valid = (j <= M.je && k <= M.ke);
  if (valid) {
    long long i = cell(M, j, k);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      sv[n] = 0.0 - (sF[n][ry][cx + 1] - sF[n][ry][cx]);
      sv[n] = sv[n] - (sG[n][ry + 1][cx] - sG[n][ry][cx]);
      M.s[n][i] = sv[n];
    }
    // the reference clears all of s at step start (src/step.f:41)
    for (int n = 4; n < M.nv; ++n) M.s[n][i] = 0.0;
  }
  if (do_norms) norm_tile(M, T, j, k, valid, sv, &red[0][0]);
}

// second stage, one block per instance: the tile partials in a fixed order
// (strided per-thread sums, then an LDS tree), the norms, the history slot
// and the "norm" status (src/step.f:104-110)
__device__ void rbc_segment(const DevMesh& M, const DevBC& b);

// with rbc = 1 the residual boundary conditions of the instance follow in
// the same workgroup (rbc, src/step.f:133-136), unless the norm stopped it
__global__ __launch_bounds__(RES_THREADS) void k_norms_final(const DevMesh* __restrict__ Ms, int rbc) {
  __shared__ double red[4][RES_THREADS];
  __shared__ int stop_s;
  const DevMesh& M = Ms[blockIdx.x];
  if (*M.status) return;
  const int tid = threadIdx.x;
  const int jd = M.jd, kd = M.kd;
  {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 1e300;
    const double* P = M.partials;
    for (int b = tid; b < M.res_nblocks; b += RES_THREADS) {
      a0 = a0 + P[4 * b + 0];
      a1 = a1 + P[4 * b + 1];
      double x = P[4 * b + 2], kb = P[4 * b + 3];
      if (x > a2 || (x == a2 && kb < a3)) { a2 = x; a3 = kb; }
    }
    red[0][tid] = a0; red[1][tid] = a1; red[2][tid] = a2; red[3][tid] = a3;
  }
  __syncthreads();
  for (int off = RES_THREADS / 2; off > 0; off >>= 1) {
    if (tid < off) {
      red[0][tid] = red[0][tid] + red[0][tid + off];
      red[1][tid] = red[1][tid] + red[1][tid + off];
      double a = red[2][tid], b = red[2][tid + off];
      double ka = red[3][tid], kb = red[3][tid + off];
      if (b > a || (b == a && kb < ka)) { red[2][tid] = b; red[3][tid] = kb; }
    }
    __syncthreads();
  }
  if (tid == 0) {
    double rsum = red[0][0], resrho = red[1][0], resmax = red[2][0], kbest = red[3][0];
    double volum = (double)(jd * kd);
    resrho = sqrt(resrho / volum);
    rsum = sqrt(rsum / volum);
    int jm = 2, km = 2, nm = 1;
    if (resmax > 0.0) {
      long long kk = (long long)kbest;
      jm = (int)(kk % jd) + 1;
      long long rest = kk / jd;
      nm = (int)(rest % 4) + 1;
      km = (int)(rest / 4) + 1;
    }
    double* N = M.norms;
    N[0] = rsum; N[1] = resrho; N[2] = resmax; N[3] = jm; N[4] = km; N[5] = nm;
    if (M.hist && M.stepc) {
      int sc = *M.stepc;
      if (sc >= 1 && sc <= M.hist_cap) M.hist[(long long)(sc - 1) * M.hist_stride + M.hist_slot] = rsum;
    }
    if (rsum > 1000.0) atomicCAS(M.status, 0, 2);
    stop_s = rsum > 1000.0;
  }
  if (!rbc) return;
  __syncthreads();
  if (stop_s) return;
  for (int ib = 0; ib < M.nbc; ++ib) {
    rbc_segment(M, M.bc[ib]);
    __syncthreads();
  }
}

// residual norms over the final residual when viscous / time-spectral terms
// were added after k_residual
__global__ __launch_bounds__(RES_THREADS) void k_norms(const DevMesh* __restrict__ Ms,
                                                       const ResTile* __restrict__ tiles) {
  __shared__ double red[4][RES_THREADS];
  const ResTile T = tiles[blockIdx.x];
  const DevMesh& M = Ms[T.mesh];
  const int tid = threadIdx.x;
  int j = T.j0 + tid % RES_TJ, k = T.k0 + tid / RES_TJ;
  bool valid = (j <= M.je && k <= M.ke);
  double sv[4] = {0.0, 0.0, 0.0, 0.0};
  if (valid) {
    long long i = cell(M, j, k);
    for (int n = 0; n < 4; ++n) sv[n] = M.s[n][i];
  }
  norm_tile(M, T, j, k, valid, sv, &red[0][0]);
}

// ===================================================== time-spectral coupling
// src/sources.f:2-27: s(nsp) -= rotf*Ds(nsp,nsp1)*q(nsp1) over interior cells
__global__ __launch_bounds__(256) void k_tsrhs(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  int nj = M.je - M.js + 1, nk = M.ke - M.ks + 1;
  long long ncell = (long long)nj * nk;
  int N = M.nspec, nsp = M.nsp;
  for (long long w = blockIdx.x * 256ll + threadIdx.x; w < ncell; w += 256ll * gridDim.x) {
    int j = M.js + (int)(w % nj), k = M.ks + (int)(w / nj);
    long long i = cell(M, j, k);
    for (int n = 0; n < M.nv; ++n) {
      double s = M.s[n][i];
      for (int nsp1 = 0; nsp1 < N; ++nsp1) {
        double dsv = M.Ds[nsp + (long long)nsp1 * N];
        s = s - M.c.rotf * dsv * M.qall[n][i + nsp1 * M.inst_stride];
      }
      M.s[n][i] = s;
    }
  }
}

// ================================================================ residual BC
// rbc (src/bc.f:88-177): zero for wall/far-field types, wake copy
// (rbcwake 1576-1633), extrapolation (2029-2094), symmetry (2160-2222),
// periodic (270-357).
__device__ void rbc_segment(const DevMesh& M, const DevBC& b) {
  double* qj = M.q[M.nq - 1];
  double* ss[4] = {M.s[0], M.s[1], M.s[2], M.s[3]};
  int js = b.js, je = b.je, ks = b.ks, ke = b.ke, idir = b.idir;
  int iadd = idir >= 0 ? 1 : -1, iadir = idir >= 0 ? idir : -idir;
  int t = threadIdx.x, nt = blockDim.x;
  int ty = b.type;
  int nj = je - js + 1, nk = ke - ks + 1;
  if (ty == 3 || ty == 4 || ty == 5 || ty == 46 || ty == 47) {
    for (int w = t; w < nj * nk; w += nt) {
      long long i = cell(M, js + w % nj, ks + w / nj);
      for (int n = 0; n < 4; ++n) ss[n][i] = 0.0;
    }
  } else if (ty == 10) {
    double foso = 0.0;
    if (iadir == 1) {
      int j = (idir == 1) ? je : js, j1 = j + iadd, j2 = j1 + iadd;
      for (int k = ks + t; k <= ke; k += nt) {
        long long i = cell(M, j, k), i1 = cell(M, j1, k), i2 = cell(M, j2, k);
        double s1 = qj[i1] / qj[i], s2 = qj[i2] / qj[i];
        for (int n = 0; n < 4; ++n) ss[n][i] = (1. + foso) * ss[n][i1] * s1 - foso * ss[n][i2] * s2;
      }
    } else if (iadir == 2) {
      int k = (idir == 2) ? ke : ks, k1 = k + iadd, k2 = k1 + iadd;
      for (int j = js + t; j <= je; j += nt) {
        long long i = cell(M, j, k), i1 = cell(M, j, k1), i2 = cell(M, j, k2);
        double s1 = qj[i1] / qj[i], s2 = qj[i2] / qj[i];
        for (int n = 0; n < 4; ++n) ss[n][i] = (1. + foso) * ss[n][i1] * s1 - foso * ss[n][i2] * s2;
      }
    }
  } else if (ty == 11) {
    if (iadir == 1) {
      int j = (idir == 1) ? je : js, j1 = (idir == 1) ? je + 1 : js - 1;
      for (int k = ks + t; k <= ke; k += nt) {
        long long i = cell(M, j, k), i1 = cell(M, j1, k);
        double sc = qj[i1] / qj[i];
        ss[0][i] = ss[0][i1] * sc; ss[1][i] = -ss[1][i1] * sc;
        ss[2][i] = ss[2][i1] * sc; ss[3][i] = ss[3][i1] * sc;
      }
    } else if (iadir == 2) {
      int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
      for (int j = js + t; j <= je; j += nt) {
        long long i = cell(M, j, k), i1 = cell(M, j, k1);
        double sc = qj[i1] / qj[i];
        ss[0][i] = ss[0][i1] * sc; ss[1][i] = ss[1][i1] * sc;
        ss[2][i] = -ss[2][i1] * sc; ss[3][i] = ss[3][i1] * sc;
      }
    }
  } else if (ty == 22) {
    for (int w = t; w < nj * nk; w += nt) {
      int j = js + w % nj, k = ks + w / nj;
      long long i = cell(M, j, k), i1;
      if (iadir == 1) i1 = cell(M, periodic_src(idir, js, je, ks, ke, M.jd, M.kd, j), k);
      else i1 = cell(M, j, periodic_src(idir, js, je, ks, ke, M.jd, M.kd, k));
      double sc = qj[i1] / qj[i];
      for (int n = 0; n < 4; ++n) ss[n][i] = ss[n][i1] * sc;
    }
  } else if (ty == 51 && iadir == 2) {
    int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
    for (int j = js + t; j <= je; j += nt) {
      int jj = M.jd - j + 1;
      long long a = cell(M, j, k), bb = cell(M, jj, k1), cc = cell(M, jj, k), d = cell(M, j, k1);
      double scale1 = qj[bb] / qj[a];
      double scale2 = qj[d] / qj[cc];
      double x[4], y[4];
      for (int n = 0; n < 4; ++n) { x[n] = ss[n][bb] * scale1; y[n] = ss[n][d] * scale2; }
      for (int n = 0; n < 4; ++n) { ss[n][a] = x[n]; ss[n][cc] = y[n]; }
    }
  }
}

__global__ __launch_bounds__(BC_THREADS) void k_rbc(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.x];
  if (*M.status) return;
  for (int ib = 0; ib < M.nbc; ++ib) {
    rbc_segment(M, M.bc[ib]);
    __syncthreads();
  }
}

// ==================================================================== LU-SGS
// Per-cell quantities ilu2d derives from the primitive state
// (src/lhs.f:524-564): u, v, uv2, ge, contravariant velocities, spectral
// radii, viscous term and the diagonal d.
template <typename T>
struct LUCoef {
  T uu, vv, uv2, ge;   // primitive velocity, kinetic energy, gamma*e - gm1*uv2
  T ri1, ri2, qq;      // metric pair and contravariant velocity of the direction
  T chk, spec;         // upwind switch and spectral term of the direction/sweep
};

struct LUCellF64 {
  double uu, vv, uv2, ge, qx, qy, cx, cy, vn, d, qqx_s, qqy_s;
};

__device__ __forceinline__ LUCellF64 lu_cell(const DevConst& C, double q1, double q2, double q3, double q4,
                                             double qj, double xx, double xy, double yx, double yy, double ug,
                                             double vg, double turmu, double ts, double epsv) {
  LUCellF64 o;
  double dj = 1.0 / q1;
  double u = q2 * dj, v = q3 * dj, e = q4 * dj, rho = q1 * qj;
  double uv2 = 0.5 * (u * u + v * v);
  double cjkl = sqrt(C.ggm1 * (e - uv2));
  double qq1 = xx * u + xy * v;
  double qqxs = -xx * ug - xy * vg + qq1;
  double qqx = fabs(qqxs);
  double rr1 = sqrt(sq(xx) + sq(xy));
  double qq2 = yx * u + yy * v;
  double qqys = -yx * ug - yy * vg + qq2;
  double qqy = fabs(qqys);
  double rr2 = sqrt(sq(yx) + sq(yy));
  double vnu = 2.0 * rr2 * rr2 * (C.rmue + turmu) / (C.rey * rho);
  o.d = 1.0 / (1.0 + ts * ((qqx + qqy + cjkl * (rr1 + rr2)) * epsv + vnu));
  o.uu = u; o.vv = v; o.uv2 = uv2;
  o.qx = qq1; o.cx = cjkl * rr1; o.qy = qq2; o.cy = cjkl * rr2; o.vn = vnu;
  o.ge = C.gamma * e - C.gm1 * uv2;
  o.qqx_s = qqxs; o.qqy_s = qqys;
  return o;
}

// A(+/-) or B(+/-) operator of a neighbour applied to its s (src/lhs.f:633-651)
__device__ __forceinline__ void lu_apply(const double gm1, double chk, double spec, double ri1, double ri2,
                                         double qq, double uu, double vv, double uv2, double ge, const double* s,
                                         double* a) {
  double a4 = chk * (ri1 * s[1] + ri2 * s[2] - qq * s[0]);
  double a2 = chk * gm1 * (uv2 * s[0] - (uu * s[1] + vv * s[2]) + s[3]);
  a[0] = a4 + spec * s[0];
  a[1] = ri1 * a2 + uu * a4 + spec * s[1];
  a[2] = ri2 * a2 + vv * a4 + spec * s[2];
  a[3] = qq * a2 + ge * a4 + spec * s[3];
}

struct LUNb {           // what a consumer needs of a neighbour cell
  double s[4];
  double ri1, ri2, qq, uu, vv, uv2, ge, chk, spec;
};

struct CellLoad {
  double q1, q2, q3, q4, qj, xx, xy, yx, yy, ug, vg, tm, ts;
};

__device__ __forceinline__ CellLoad load_cell(const DevMesh& M, long long i) {
  CellLoad c;
  c.q1 = M.q[0][i]; c.q2 = M.q[1][i]; c.q3 = M.q[2][i]; c.q4 = M.q[3][i]; c.qj = M.q[M.nq - 1][i];
  c.xx = M.xx[i]; c.xy = M.xy[i]; c.yx = M.yx[i]; c.yy = M.yy[i];
  c.ug = M.has_gv ? M.ug[i] : 0.0;
  c.vg = M.has_gv ? M.vg[i] : 0.0;
  c.tm = M.turmu[i];
  c.ts = M.tscale[i];
  return c;
}

__device__ __forceinline__ LUCellF64 lu_of(const DevConst& C, const CellLoad& c, double epsv) {
  return lu_cell(C, c.q1, c.q2, c.q3, c.q4, c.qj, c.xx, c.xy, c.yx, c.yy, c.ug, c.vg, c.tm, c.ts, epsv);
}

// neighbour record for the j direction (A term) of sweep direction dir (+1 fwd, -1 bwd)
__device__ __forceinline__ void nb_j(const CellLoad& c, const LUCellF64& L, double eps2, int dir, LUNb& o) {
  double qqx = L.qqx_s;
  double sp1 = fabs(qqx) + L.cx;
  double sm2 = eps2 * sp1;
  if (dir > 0) {
    o.chk = 0.5 + fsign(0.5, qqx + L.cx);
    o.spec = o.chk * (qqx + sp1) + sm2;
  } else {
    o.chk = 0.5 - fsign(0.5, qqx - L.cx);
    o.spec = o.chk * (qqx - sp1) - sm2;
  }
  o.ri1 = c.xx; o.ri2 = c.xy; o.qq = L.qx; o.uu = L.uu; o.vv = L.vv; o.uv2 = L.uv2; o.ge = L.ge;
}

// neighbour record for the k direction (B term)
__device__ __forceinline__ void nb_k(const CellLoad& c, const LUCellF64& L, double eps2, int dir, LUNb& o) {
  double qqy = L.qqy_s;
  double sp1 = fabs(qqy) + L.cy;
  double sm2 = eps2 * sp1;
  if (dir > 0) {
    o.chk = 0.5 + fsign(0.5, qqy + L.cy);
    o.spec = o.chk * (qqy + sp1) + L.vn + sm2;
  } else {
    o.chk = 0.5 - fsign(0.5, qqy - L.cy);
    o.spec = o.chk * (qqy - sp1) - L.vn - sm2;
  }
  o.ri1 = c.yx; o.ri2 = c.yy; o.qq = L.qy; o.uu = L.uu; o.vv = L.vv; o.uv2 = L.uv2; o.ge = L.ge;
}

// ----------------------------------------------------------------------
// LU-SGS in three kernels.
//
//  k_lu_prep    every cell in parallel: the quantities ilu2d derives per cell
//               (src/lhs.f:524-564), s*tscale (src/step.f:141-148) and the
//               point-Jacobi update of the halo sides (src/lhs.f:568-598),
//               stored as 20-value records in the order the sweeps consume
//               them (layout below).
//  k_lu_sweep   one wavefront per instance and direction.  Lane l owns RL
//               consecutive rows k = ks + l*RL + r and walks along j, one
//               column per step, lane l one step behind lane l-1 (forward):
//               cell (j,k) is done at step t = (j-js) + l.  Its j neighbour is
//               the same lane's previous step (registers); its k neighbour is
//               either the lane's own row below (same step) or lane l-1's top
//               row of the previous step, passed by a DPP wavefront shift.
//               No barrier and no LDS round trip sits on the dependent
//               chain; records stream in D steps ahead.  The backward sweep
//               mirrors this (t = (je-j) + (NL-1-l), DPP shift the other way).
//               The result of every cell is the reference's (src/lhs.f:
//               616-714 forward, 719-800 backward): a hyperplane order and
//               this skewed order evaluate each cell from the same operands.
//  k_lu_finish  every cell in parallel: conservative restore of the in-place
//               primitive conversion (src/lhs.f:805-811), the reference's s
//               after ilu2d, and optionally the q update (src/step.f:205-211)
//               and monitor (src/step.f:353-385).
//
// Record (20 values): s0..s3, sv, di, uu, vv, uv2, ge, xx, xy, qx, cj, pj,
// yx, yy, qy, ck, pk (cj/pj and ck/pk: switch and spectral term of the
// j and k directions for the sweep that reads the record).  FP32 packs it in
// 5 float4, FP64 in 10 double2; vector q of entry n sits at plane q.
//
// Workspace of one instance, in 16-byte vectors (NVEC = 10 planes reserved
// so both precisions fit):
//   F   forward records   entry n = (t*RL + r)*64 + l, t = (j-js) + l
//   B   backward records  entry n = (t*RL + r)*64 + l, t = (je-j) + (NL-1-l);
//       its s vectors receive the forward result, then the backward result
//   HCF, HCB  halo column j = js-1 / je+1, entry r*64 + l
//   HRF, HRB  halo row    k = ks-1 / ke+1, entry j-js
//
// T is the arithmetic of the sweep: double reproduces the reference, float
// is the mixed-precision defect-correction variant (records rounded from the
// FP64 values computed in k_lu_prep).
// ----------------------------------------------------------------------

constexpr int LU_NVEC = 10;        // 16-byte planes reserved per array
constexpr int LU32_R = 32;         // rows of one FP32 sweep workgroup (two lanes per row)
constexpr int LU_MAXRL = 8;        // rows per lane: at most 512 rows

template <typename T> struct LUV;
template <> struct LUV<float> { static constexpr int N = 5; };
template <> struct LUV<double> { static constexpr int N = 10; };

struct LUGeom {
  int ni, nk, RL, NL, T;           // FP64 sweep: NL lanes, T steps
  int W, TW, TA;                   // FP32 sweep: W workgroups of LU32_R rows, TW steps each,
                                   // TA >= TW step rows of the record arrays (a multiple of 8)
  long long nrec;                  // FP64: entries of F and of B
  long long off_b, off_hcf, off_hcb, off_hrf, off_hrb;   // FP64 arrays (16-byte vectors)
  long long ncell;                 // FP32: interior cells
  long long off_f32, off_b32, off_cb, off_h, off_h2;   // FP32 arrays
  long long off_tk;                // FP32 workgroup tickets (forward, backward): one vector
  long long total;                 // vectors the workspace must hold (either layout)
};

__host__ __device__ inline LUGeom lu_geom(int ni, int nk, int RL) {
  LUGeom g;
  g.ni = ni; g.nk = nk; g.RL = RL;
  g.NL = (nk + RL - 1) / RL;
  g.T = ni + g.NL - 1;
  g.W = (nk + LU32_R - 1) / LU32_R;
  g.TW = ni + LU32_R - 1;
  g.TA = (g.TW + 7) & ~7;
  g.nrec = (long long)g.T * RL * 64;
  long long hc = (long long)RL * 64;
  g.off_b = LU_NVEC * g.nrec;
  g.off_hcf = 2 * LU_NVEC * g.nrec;
  g.off_hcb = g.off_hcf + LU_NVEC * hc;
  g.off_hrf = g.off_hcb + LU_NVEC * hc;
  g.off_hrb = g.off_hrf + (long long)LU_NVEC * ni;
  long long t64 = g.off_hrb + (long long)LU_NVEC * ni;
  // FP32, all in step layout (per block and sweep step, a row of 64 lanes):
  // the forward records (4 vectors of 16 bytes per lane + one 8-byte pair:
  // 288 16-byte units), the backward records (4 vectors: 256 units), cB (8
  // bytes per lane: 32 units), the forward and the backward results (16
  // bytes per lane, each value beside the sweep's tag: 64 units each; the
  // edge rows' results are also the inter-workgroup mail); then the two
  // ticket counters
  g.ncell = (long long)ni * nk;
  g.off_f32 = 0;
  g.off_b32 = g.off_f32 + 288ll * g.W * g.TA;
  g.off_cb = g.off_b32 + 256ll * g.W * g.TA;
  g.off_h = g.off_cb + 32ll * g.W * g.TW;
  g.off_h2 = g.off_h + 64ll * g.W * g.TW;
  g.off_tk = g.off_h2 + 64ll * g.W * g.TW;
  long long t32 = g.off_tk + 1;
  g.total = t64 > t32 ? t64 : t32;
  return g;
}

__device__ __forceinline__ LUGeom lu_geom(const DevMesh& M) {
  return lu_geom(M.je - M.js + 1, M.ke - M.ks + 1, M.lu_rl);
}

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
struct D2 { double x, y; };
struct F4 { float x, y, z, w; };
struct F2 { float x, y; };

// buffer descriptor built from wave-uniform values (keeps it in SGPRs, so
// buffer accesses need no waterfall loop)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t uniform_rsrc(void* p, long long bytes) {
  unsigned long long v = (unsigned long long)p;
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  void* up = (void*)(((unsigned long long)hi << 32) | lo);
  int nb = __builtin_amdgcn_readfirstlane((int)bytes);
  return __builtin_amdgcn_make_buffer_rsrc(up, 0, nb, 0x00020000);
}

// one array of the workspace: plane q holds vector q of every entry
struct LUArr {
  D2* base;
  long long n;     // entries per plane
  __device__ __forceinline__ D2* at(int q, long long e) const { return base + q * n + e; }
};

__device__ __forceinline__ LUArr lu_arr(const DevMesh& M, long long off, long long n) {
  LUArr a;
  a.base = reinterpret_cast<D2*>(M.lu) + off;
  a.n = n;
  return a;
}

// write one record (FP64 values, rounded for the FP32 layout)
template <typename T>
__device__ __forceinline__ void rec_put(const LUArr& a, long long e, const double* f) {
  if (sizeof(T) == 8) {
#pragma unroll
    for (int q = 0; q < 10; ++q) *a.at(q, e) = D2{f[2 * q], f[2 * q + 1]};
  } else {
#pragma unroll
    for (int q = 0; q < 5; ++q)
      *reinterpret_cast<F4*>(a.at(q, e)) = F4{(float)f[4 * q], (float)f[4 * q + 1], (float)f[4 * q + 2],
                                              (float)f[4 * q + 3]};
  }
}

template <typename T>
__device__ __forceinline__ void rec_get(const LUArr& a, long long e, T* f) {
  if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int q = 0; q < 10; ++q) { D2 v = *a.at(q, e); f[2 * q] = v.x; f[2 * q + 1] = v.y; }
  } else {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      F4 v = *reinterpret_cast<const F4*>(a.at(q, e));
      f[4 * q] = v.x; f[4 * q + 1] = v.y; f[4 * q + 2] = v.z; f[4 * q + 3] = v.w;
    }
  }
}

template <typename T>
__device__ __forceinline__ void rec_unpack(const u32x4_t* v, T* f) {
  if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int q = 0; q < 10; ++q) { D2 d = __builtin_bit_cast(D2, v[q]); f[2 * q] = d.x; f[2 * q + 1] = d.y; }
  } else {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      F4 d = __builtin_bit_cast(F4, v[q]);
      f[4 * q] = d.x; f[4 * q + 1] = d.y; f[4 * q + 2] = d.z; f[4 * q + 3] = d.w;
    }
  }
}

// entry of interior cell (j,k) in F (fwd) or B (bwd)
__device__ __forceinline__ long long lu_entry(const DevMesh& M, const LUGeom& g, int j, int k, int dir) {
  int kk = k - M.ks, l = kk / g.RL, r = kk - l * g.RL;
  int t = dir > 0 ? (j - M.js) + l : (M.je - j) + (g.NL - 1 - l);
  return ((long long)t * g.RL + r) * 64 + l;
}

// FP32 results of interior cell (j,k): workgroup w = (k-ks)/32 owns rows
// ks+32w .. ks+32w+31, row r = (k-ks)%32 on lanes 2r (components 1, 2) and
// 2r+1 (3, 4); forward step t = (j-js) + r, backward step t = (je-j) +
// (31-r).  Returns the first of the two 16-byte units holding the cell's
// four components in a result array (16 bytes per lane: component, tag,
// component, tag; 64 units per block and step).
__device__ __forceinline__ long long lu_entry32(const DevMesh& M, const LUGeom& g, int j, int k, int dir) {
  int kk = k - M.ks, w = kk / LU32_R, r = kk % LU32_R;
  int t = dir > 0 ? (j - M.js) + r : (M.je - j) + (LU32_R - 1 - r);
  return ((long long)w * g.TW + t) * 64 + 2 * r;
}

__global__ __launch_bounds__(256) void k_lu_prep(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const DevConst& C = M.c;
  const int jd = M.jd, kd = M.kd, js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const double eps2 = C.epse * 2.;
  const double epsv = 1. + eps2;
  const LUGeom g = lu_geom(M);
  const LUArr F = lu_arr(M, 0, g.nrec), B = lu_arr(M, g.off_b, g.nrec);
  const LUArr HCF = lu_arr(M, g.off_hcf, (long long)g.RL * 64), HCB = lu_arr(M, g.off_hcb, (long long)g.RL * 64);
  const LUArr HRF = lu_arr(M, g.off_hrf, g.ni), HRB = lu_arr(M, g.off_hrb, g.ni);
  long long n2 = (long long)jd * kd;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    int j = 1 + (int)(i % jd), k = 1 + (int)(i / jd);
    double ts = M.tscale[i];
    bool jin = (j >= js && j <= je), kin = (k >= ks && k <= ke);
    if (!jin && !kin) {  // corner: only s*tscale
      for (int n = 0; n < 4; ++n) M.s[n][i] = M.s[n][i] * ts;
      continue;
    }
    CellLoad c = load_cell(M, i);
    LUCellF64 L = lu_of(C, c, epsv);
    bool interior = jin && kin;
    double st[4];
    for (int n = 0; n < 4; ++n) {
      st[n] = M.s[n][i] * ts;
      if (!interior) {
        st[n] = st[n] * L.d;
        M.s[n][i] = st[n];
      }
    }
    LUNb jf, jb, kf, kb;
    nb_j(c, L, eps2, +1, jf);
    nb_j(c, L, eps2, -1, jb);
    nb_k(c, L, eps2, +1, kf);
    nb_k(c, L, eps2, -1, kb);
    double f[20] = {st[0], st[1], st[2], st[3], ts * 0.5, L.d, L.uu, L.vv, L.uv2, L.ge,
                    c.xx, c.xy, L.qx, jf.chk, jf.spec, c.yx, c.yy, L.qy, kf.chk, kf.spec};
    double b[20];
    for (int n = 0; n < 20; ++n) b[n] = f[n];
    b[13] = jb.chk; b[14] = jb.spec; b[18] = kb.chk; b[19] = kb.spec;
    if (interior) {
      rec_put<double>(F, lu_entry(M, g, j, k, +1), f);
      rec_put<double>(B, lu_entry(M, g, j, k, -1), b);
    } else if (kin) {  // halo column
      int kk = k - ks, l = kk / g.RL, r = kk - l * g.RL;
      long long e = (long long)r * 64 + l;
      if (j == js - 1) rec_put<double>(HCF, e, f);
      if (j == je + 1) rec_put<double>(HCB, e, b);
    } else {           // halo row
      long long e = j - js;
      if (k == ks - 1) rec_put<double>(HRF, e, f);
      if (k == ke + 1) rec_put<double>(HRB, e, b);
    }
  }
}

// what a cell needs of a neighbour: its new s and the static terms of the
// direction it is a neighbour in
template <typename T>
struct NbT {
  T s[4];
  T uu, vv, uv2, ge, ri1, ri2, qq, chk, spec;
};

// A(+/-) or B(+/-) applied to the neighbour's s (src/lhs.f:633-651, 665-681)
template <typename T>
__device__ __forceinline__ void apply_nb(T gm1, const NbT<T>& n, T* a) {
  T a4 = n.chk * (n.ri1 * n.s[1] + n.ri2 * n.s[2] - n.qq * n.s[0]);
  T a2 = n.chk * gm1 * (n.uv2 * n.s[0] - (n.uu * n.s[1] + n.vv * n.s[2]) + n.s[3]);
  a[0] = a4 + n.spec * n.s[0];
  a[1] = n.ri1 * a2 + n.uu * a4 + n.spec * n.s[1];
  a[2] = n.ri2 * a2 + n.vv * a4 + n.spec * n.s[2];
  a[3] = n.qq * a2 + n.ge * a4 + n.spec * n.s[3];
}

template <typename T>
__device__ __forceinline__ void jrole(const T* s, const T* f, NbT<T>& n) {
  n.s[0] = s[0]; n.s[1] = s[1]; n.s[2] = s[2]; n.s[3] = s[3];
  n.uu = f[6]; n.vv = f[7]; n.uv2 = f[8]; n.ge = f[9];
  n.ri1 = f[10]; n.ri2 = f[11]; n.qq = f[12]; n.chk = f[13]; n.spec = f[14];
}
template <typename T>
__device__ __forceinline__ void krole(const T* s, const T* f, NbT<T>& n) {
  n.s[0] = s[0]; n.s[1] = s[1]; n.s[2] = s[2]; n.s[3] = s[3];
  n.uu = f[6]; n.vv = f[7]; n.uv2 = f[8]; n.ge = f[9];
  n.ri1 = f[15]; n.ri2 = f[16]; n.qq = f[17]; n.chk = f[18]; n.spec = f[19];
}

// wavefront shift of one value: DIR > 0 lane i receives lane i-1
// (wave_shr:1), DIR < 0 lane i receives lane i+1 (wave_shl:1)
template <int DIR>
__device__ __forceinline__ int dpp_shift(int v) {
  return __builtin_amdgcn_update_dpp(0, v, DIR > 0 ? 0x138 : 0x130, 0xf, 0xf, false);
}
template <int DIR>
__device__ __forceinline__ float dpp_shift(float v) {
  return __builtin_bit_cast(float, dpp_shift<DIR>(__builtin_bit_cast(int, v)));
}
template <int DIR>
__device__ __forceinline__ double dpp_shift(double v) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = dpp_shift<DIR>((int)(b & 0xffffffffll));
  int hi = dpp_shift<DIR>((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <int DIR, typename T>
__device__ __forceinline__ NbT<T> dpp_shift_nb(const NbT<T>& x) {
  NbT<T> y;
#pragma unroll
  for (int v = 0; v < 4; ++v) y.s[v] = dpp_shift<DIR>(x.s[v]);
  y.uu = dpp_shift<DIR>(x.uu); y.vv = dpp_shift<DIR>(x.vv); y.uv2 = dpp_shift<DIR>(x.uv2);
  y.ge = dpp_shift<DIR>(x.ge); y.ri1 = dpp_shift<DIR>(x.ri1); y.ri2 = dpp_shift<DIR>(x.ri2);
  y.qq = dpp_shift<DIR>(x.qq); y.chk = dpp_shift<DIR>(x.chk); y.spec = dpp_shift<DIR>(x.spec);
  return y;
}

constexpr int LU_HR = 13;   // values of a halo-row record kept in LDS

template <typename T>
__host__ __device__ inline size_t lu_lds_bytes(int ni) {
  return sizeof(T) * (size_t)LU_HR * (size_t)ni;
}

template <typename T, int DIR, int RL, int D>
__global__ __launch_bounds__(64) void k_lu_sweep(const DevMesh* __restrict__ Ms) {
  extern __shared__ __attribute__((aligned(16))) double lds_raw[];
  constexpr int NV = LUV<T>::N;
  const DevMesh& M = Ms[blockIdx.x];
  if (*M.status) return;
  const int js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const LUGeom g = lu_geom(M);
  const int l = threadIdx.x;
  const T gm1 = (T)M.c.gm1;
  const int NT = g.T;
  // records this sweep reads, and where its results go (the s vectors of B:
  // forward writes entry T-1-t, backward overwrites entry t)
  const long long roff = DIR > 0 ? 0 : g.off_b;
  __amdgpu_buffer_rsrc_t rec[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) rec[q] = uniform_rsrc(reinterpret_cast<D2*>(M.lu) + roff + q * g.nrec, g.nrec * 16);
  __amdgpu_buffer_rsrc_t outq[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) outq[q] = uniform_rsrc(reinterpret_cast<D2*>(M.lu) + g.off_b + q * g.nrec, g.nrec * 16);
  // halo row (k neighbour of the boundary lane) staged in LDS
  T* hr = reinterpret_cast<T*>(lds_raw);
  {
    const LUArr HR = lu_arr(M, DIR > 0 ? g.off_hrf : g.off_hrb, g.ni);
    for (int w = l; w < g.ni; w += 64) {
      T f[20];
      rec_get<T>(HR, w, f);
      NbT<T> n;
      krole(f, f, n);
      T* x = hr + LU_HR * w;
      x[0] = n.s[0]; x[1] = n.s[1]; x[2] = n.s[2]; x[3] = n.s[3]; x[4] = n.uu; x[5] = n.vv; x[6] = n.uv2;
      x[7] = n.ge; x[8] = n.ri1; x[9] = n.ri2; x[10] = n.qq; x[11] = n.chk; x[12] = n.spec;
    }
  }
  __syncthreads();
  // j neighbour of every row's first cell: the halo column
  NbT<T> A[RL];
  {
    const LUArr HC = lu_arr(M, DIR > 0 ? g.off_hcf : g.off_hcb, (long long)g.RL * 64);
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      T f[20];
      rec_get<T>(HC, (long long)r * 64 + l, f);
      jrole(f, f, A[r]);
    }
  }
  NbT<T> kout;
  jrole(hr, hr, kout);   // any finite values: only read by lanes that are inactive
  const bool blane = DIR > 0 ? (l == 0) : (l == g.NL - 1);
  const int vo = l * 16;
  u32x4_t P[D][RL][NV];
  auto fetch = [&](int tp, u32x4_t (&dst)[RL][NV]) {
#pragma unroll
    for (int r = 0; r < RL; ++r)
#pragma unroll
      for (int q = 0; q < NV; ++q)
        dst[r][q] = __builtin_amdgcn_raw_buffer_load_b128(rec[q], vo, (tp * RL + r) * 1024, 0);
  };
#pragma unroll
  for (int d = 0; d < D - 1; ++d)
    if (d < NT) fetch(d, P[d]);
  for (int t0 = 0; t0 < NT; t0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int t = t0 + d;
      if (t >= NT) break;
      if (t + D - 1 < NT) fetch(t + D - 1, P[(d + D - 1) % D]);
      const int j = DIR > 0 ? js + t - l : je - t + (g.NL - 1 - l);
      const bool lact = l < g.NL && j >= js && j <= je;
      NbT<T> kin = dpp_shift_nb<DIR>(kout);
      if (blane) {
        int jh = j - js;
        jh = jh < 0 ? 0 : (jh >= g.ni ? g.ni - 1 : jh);
        const T* x = hr + LU_HR * jh;
        kin.s[0] = x[0]; kin.s[1] = x[1]; kin.s[2] = x[2]; kin.s[3] = x[3];
        kin.uu = x[4]; kin.vv = x[5]; kin.uv2 = x[6]; kin.ge = x[7]; kin.ri1 = x[8]; kin.ri2 = x[9];
        kin.qq = x[10]; kin.chk = x[11]; kin.spec = x[12];
      }
      const int ostep = DIR > 0 ? NT - 1 - t : t;
#pragma unroll
      for (int rr = 0; rr < RL; ++rr) {
        const int r = DIR > 0 ? rr : RL - 1 - rr;
        const int k = ks + l * RL + r;
        const bool v = lact && k <= ke;
        T f[20];
        rec_unpack<T>(P[d][r], f);
        T a[4], b[4], sn[4];
        apply_nb(gm1, A[r], a);
        apply_nb(gm1, kin, b);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          if (DIR > 0) sn[w] = (f[w] + f[4] * (a[w] + b[w])) * f[5];
          else sn[w] = f[w] - f[4] * (a[w] + b[w]) * f[5];
        }
        if (v) {
          const int so = (ostep * RL + r) * 1024;
          if constexpr (sizeof(T) == 8) {
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, D2{sn[0], sn[1]}), outq[0], vo, so, 0);
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, D2{sn[2], sn[3]}), outq[1], vo, so, 0);
          } else {
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, F4{sn[0], sn[1], sn[2], sn[3]}),
                                                   outq[0], vo, so, 0);
          }
          jrole(sn, f, A[r]);
          krole(sn, f, kin);
        }
      }
      kout = kin;
    }
  }
}

// ------------------------------------------------- FP32 LU-SGS (matrix form)
// The mixed-precision sweeps evaluate the same update as ilu2d with the
// operator assembled per cell in FP64 and rounded once (k_lu_prep32):
//   forward   s(j,k) = u + PA . s(j-1,k) + PB . s(j,k-1)
//             u = st*di,  PA = sv*di*A+(j-1,k),  PB = sv*di*B+(j,k-1)
//   backward  s(j,k) = s_fwd(j,k) + PA' . s(j+1,k) + PB' . s(j,k+1)
//             PA' = -sv*di*A-(j+1,k),  PB' = -sv*di*B-(j,k+1)
// (src/lhs.f:616-714 and 719-800 with a(j) = A . s, b(j) = B . s written
// as 4x4 matrices).  Neighbours in the halo ring hold fixed point-Jacobi
// values (src/lhs.f:568-598), so their terms are folded into u (forward)
// and into cB (backward, added by the forward sweep to its output).
//
// A workgroup sweeps a block of LU32_R = 32 rows.  Lane l = 2r + h holds
// row r and components 2h, 2h+1 ("its pair"); lane l^1 holds the other
// pair of the same row.  Lane l's view of a row's four components is
// (its pair, the other pair), so the records store each matrix row's
// columns in that order and no lane needs a select.  Per lane and sweep
// step:
//   forward record  4 rows of 4 floats (PA rows 2h, 2h+1; PB rows 2h,
//                   2h+1) and the pair of u
//   backward record 4 rows (PA' rows, PB' rows); cB and the forward result
//                   as pairs
// all stored by forward step: for block w and step t the forward record is
// 16-byte units ((w*TA + t)*288 + q*64 + l) for rows q = 0..3 and bytes
// 4096 + 8l after the record's start for u, so a producer reads a step as
// four coalesced 1 KiB rows and one 512-byte row; the backward record (256
// units) is kept at the cell's forward step too.
//   H  (16 bytes per lane and step: the pair, each value beside the
//      sweep's tag): s_fwd, written by the forward sweep
//   H2 (the same): the result, written by the backward sweep
// (the edge row's lanes of H and H2 are also the inter-workgroup mail)
// The halo-ring s values themselves are scaled in k_lu_finish (the prep
// reads the unscaled neighbours).

// cache-policy operand (sc0 and sc1 bits: system scope) of the result
// stores of the sweeps
constexpr int LU32_SYS = 17;

// rows of the 4x4 matrix M with a = M . s for a neighbour's A(+/-) or
// B(+/-) term (src/lhs.f:633-651)
__device__ __forceinline__ void nb_matrix(const LUNb& n, double gm1, double m[16]) {
  double v4[4] = {-n.chk * n.qq, n.chk * n.ri1, n.chk * n.ri2, 0.0};
  double cg = n.chk * gm1;
  double v2[4] = {cg * n.uv2, -cg * n.uu, -cg * n.vv, cg};
  for (int i = 0; i < 4; ++i) {
    m[0 * 4 + i] = v4[i];
    m[1 * 4 + i] = n.ri1 * v2[i] + n.uu * v4[i];
    m[2 * 4 + i] = n.ri2 * v2[i] + n.vv * v4[i];
    m[3 * 4 + i] = n.qq * v2[i] + n.ge * v4[i];
  }
  for (int w = 0; w < 4; ++w) m[w * 4 + w] += n.spec;
}

// Prep tiles.  A workgroup of LP_WAVES waves covers 8 rows of a block and
// LP_WAVES*8 consecutive sweep steps; row r at step t is cell jr = t - r,
// so its entries lie in a parallelogram of cells whose bounding box, with
// the one-cell ring the neighbour terms reach, is LP_RJ x LP_RK cells.  The
// per-cell quantities of ilu2d (src/lhs.f:524-564) are formed once per cell
// of that box in LDS; each cell then assembles its operators from itself and
// its four neighbours there.
constexpr int LP_WAVES = 4, LP_STEPS = 8 * LP_WAVES;
constexpr int LP_RJ = LP_STEPS + 7 + 2, LP_RK = 8 + 2, LP_N = LP_RJ * LP_RK;

struct LUPrepLds {
  double uu[LP_N], vv[LP_N], uv2[LP_N], ge[LP_N], qx[LP_N], qy[LP_N], cx[LP_N], cy[LP_N], vn[LP_N], d[LP_N],
      qqx[LP_N], qqy[LP_N];
  double xx[LP_N], xy[LP_N], yx[LP_N], yy[LP_N], ts[LP_N];
  double sh[4][LP_N];   // s * tscale * d (src/step.f:141-148, src/lhs.f:568-598)
};
constexpr size_t LP_LDS = sizeof(LUPrepLds);

// neighbour term of a cell from the cell at LDS index n: matrix scaled by
// f, or folded into c when the neighbour is a halo-ring cell (its s after
// s*tscale and point Jacobi)
__device__ __forceinline__ void lu32_term(const LUPrepLds& S, int n, bool halo, int jdir, int dir, double f,
                                          double eps2, double gm1, double* P, double* c) {
  CellLoad cn;
  cn.xx = S.xx[n]; cn.xy = S.xy[n]; cn.yx = S.yx[n]; cn.yy = S.yy[n];
  LUCellF64 Ln;
  Ln.uu = S.uu[n]; Ln.vv = S.vv[n]; Ln.uv2 = S.uv2[n]; Ln.ge = S.ge[n]; Ln.qx = S.qx[n]; Ln.qy = S.qy[n];
  Ln.cx = S.cx[n]; Ln.cy = S.cy[n]; Ln.vn = S.vn[n]; Ln.qqx_s = S.qqx[n]; Ln.qqy_s = S.qqy[n];
  LUNb nb;
  if (jdir) nb_j(cn, Ln, eps2, dir, nb);
  else nb_k(cn, Ln, eps2, dir, nb);
  double m[16];
  nb_matrix(nb, gm1, m);
  for (int q = 0; q < 16; ++q) P[q] = f * m[q];
  if (halo) {
    double sh[4];
    for (int w = 0; w < 4; ++w) sh[w] = S.sh[w][n];
    for (int w = 0; w < 4; ++w) c[w] += P[w * 4 + 0] * sh[0] + P[w * 4 + 1] * sh[1] + P[w * 4 + 2] * sh[2] + P[w * 4 + 3] * sh[3];
    for (int q = 0; q < 16; ++q) P[q] = 0.0;
  }
}

// matrix row `row` of P with its columns in lane pair h's order: its own
// pair (2h, 2h+1) first, then the other pair
__device__ __forceinline__ F4 pair_row(const double* P, int row, int h) {
  const int o = 2 * h, p = 2 - 2 * h;
  return F4{(float)P[row * 4 + o], (float)P[row * 4 + o + 1], (float)P[row * 4 + p], (float)P[row * 4 + p + 1]};
}

__global__ __launch_bounds__(64 * LP_WAVES) void k_lu_prep32(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const double eps2 = M.c.epse * 2.;
  const double epsv = 1. + eps2;
  const LUGeom g = lu_geom(M);
  D2* lu = reinterpret_cast<D2*>(M.lu);
  constexpr int RG = LU32_R / 8;   // 8-row groups per block
  const int nthr = 64 * LP_WAVES;
  // tile of this workgroup: block w, rows 8*lb .. 8*lb+7, steps T0 .. T0+LP_STEPS-1
  const int tiles_t = (g.TA + LP_STEPS - 1) / LP_STEPS;
  if ((int)blockIdx.x >= g.W * RG * tiles_t) return;
  const int tb = (int)blockIdx.x % tiles_t, rest = (int)blockIdx.x / tiles_t;
  const int lb = rest % RG, w = rest / RG;
  const int T0 = tb * LP_STEPS, L0 = lb * 8;
  // box origin: cell (jr, kk) = (T0 - L0 - 8, LU32_R*w + L0 - 1) is LDS cell 0
  const int jr0 = T0 - L0 - 8, kk0 = LU32_R * w + L0 - 1;
  extern __shared__ __attribute__((aligned(16))) double lp_raw[];
  LUPrepLds& S = *reinterpret_cast<LUPrepLds*>(lp_raw);
  for (int n = threadIdx.x; n < LP_N; n += nthr) {
    const int jr = jr0 + n % LP_RJ, kk = kk0 + n / LP_RJ;
    if (jr < -1 || jr > g.ni || kk < -1 || kk > g.nk) continue;
    const long long i = cell(M, js + jr, ks + kk);
    const CellLoad c = load_cell(M, i);
    double sv[4];
    for (int q = 0; q < 4; ++q) sv[q] = M.s[q][i];
    const LUCellF64 L = lu_of(M.c, c, epsv);
    S.uu[n] = L.uu; S.vv[n] = L.vv; S.uv2[n] = L.uv2; S.ge[n] = L.ge; S.qx[n] = L.qx; S.qy[n] = L.qy;
    S.cx[n] = L.cx; S.cy[n] = L.cy; S.vn[n] = L.vn; S.d[n] = L.d; S.qqx[n] = L.qqx_s; S.qqy[n] = L.qqy_s;
    S.xx[n] = c.xx; S.xy[n] = c.xy; S.yx[n] = c.yx; S.yy[n] = c.yy; S.ts[n] = c.ts;
    for (int q = 0; q < 4; ++q) S.sh[q][n] = sv[q] * c.ts * L.d;
  }
  __syncthreads();
  // one thread per (step, row) cell, writing the records of the row's two
  // lanes: 8 consecutive threads are 8 consecutive rows at one step, so
  // each record row of a step is stored as whole 256-byte runs.  Cells
  // outside the grid (rows past nk, steps before a row's first cell or
  // after its last) keep zero records: the workspace is zero-filled when
  // the case is created and nothing else writes there.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = L0 + (lane & 7), t = T0 + 8 * wave + (lane >> 3);
  const int kk = LU32_R * w + r, jr = t - r;
  if (t >= g.TA || kk >= g.nk || jr < 0 || jr >= g.ni) return;
  const int j = js + jr, k = ks + kk;
  const int n0 = (kk - kk0) * LP_RJ + (jr - jr0);
  const double di = S.d[n0], f = (S.ts[n0] * 0.5) * di;
  const double gm1 = M.c.gm1;
  double u[4], cb[4] = {0.0, 0.0, 0.0, 0.0}, PA[16], PB[16], QA[16], QB[16];
  for (int q = 0; q < 4; ++q) u[q] = S.sh[q][n0];
  lu32_term(S, n0 - 1, j == js, 1, +1, f, eps2, gm1, PA, u);
  lu32_term(S, n0 - LP_RJ, k == ks, 0, +1, f, eps2, gm1, PB, u);
  lu32_term(S, n0 + 1, j == je, 1, -1, -f, eps2, gm1, QA, cb);
  lu32_term(S, n0 + LP_RJ, k == ke, 0, -1, -f, eps2, gm1, QB, cb);
  const long long rec = (long long)w * g.TA + t;
  F4* rf = reinterpret_cast<F4*>(lu + g.off_f32 + rec * 288);
  F4* rb = reinterpret_cast<F4*>(lu + g.off_b32 + rec * 256);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int l = 2 * r + h;
    rf[0 * 64 + l] = pair_row(PA, 2 * h, h);
    rf[1 * 64 + l] = pair_row(PA, 2 * h + 1, h);
    rf[2 * 64 + l] = pair_row(PB, 2 * h, h);
    rf[3 * 64 + l] = pair_row(PB, 2 * h + 1, h);
    rb[0 * 64 + l] = pair_row(QA, 2 * h, h);
    rb[1 * 64 + l] = pair_row(QA, 2 * h + 1, h);
    rb[2 * 64 + l] = pair_row(QB, 2 * h, h);
    rb[3 * 64 + l] = pair_row(QB, 2 * h + 1, h);
  }
  // u of lanes 2r, 2r+1 and cB of the cell (both lanes' pairs, one unit)
  rf[256 + r] = F4{(float)u[0], (float)u[1], (float)u[2], (float)u[3]};
  reinterpret_cast<F4*>(lu + g.off_cb)[((long long)w * g.TW + t) * 32 + r] =
      F4{(float)cb[0], (float)cb[1], (float)cb[2], (float)cb[3]};
}

// Bounded waits.  Every wait of the FP32 sweeps gives up after LU32_WAIT
// ticks of the 100 MHz real-time counter (0.5 s; a normal wait is a few
// microseconds), raising the workgroup's abort flag, so every wave of every
// workgroup always reaches the end of the kernel; the sweep then reports
// status 3.
constexpr unsigned long long LU32_WAIT = 50000000ull;

__device__ __forceinline__ unsigned long long rt_now() { return __builtin_amdgcn_s_memrealtime(); }

template <int SLEEP>
__device__ __forceinline__ bool spin_until_ge(int* w, int target, int* abort_flag) {
  if (__hip_atomic_load(w, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) return true;
  const unsigned long long t0 = rt_now();
  for (;;) {
    if (__hip_atomic_load(w, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) return true;
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
    if (rt_now() - t0 > LU32_WAIT) break;
    __builtin_amdgcn_s_sleep(SLEEP);
  }
  __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return false;
}

// Inter-workgroup mail.  Every lane stores its result of sweep step t as
// 16 bytes in its direction's result array, two 8-byte words each holding
// a component's bits in the low half and the sweep's tag in the high half,
// by one aligned 16-byte store with the system-scope cache policy.  The
// edge row's two lanes of step t (32 bytes, components 1-4) are the mail
// entry the neighbour block reads: it accepts the entry only when all four
// words, each read by one 8-byte system-scope atomic load, carry the
// running sweep's tag.  An entry left by an earlier sweep (an older tag) or
// one still being written (an older tag in some word) is therefore never
// taken, and only the sweep writes these rows: they need no reset between
// sweeps.  Loads go through global (not generic) pointers: flat
// instructions would also count against the LDS wait counter and stall the
// LDS hand-offs behind them.  System scope: writer and reader may sit on
// different XCDs.
typedef __attribute__((address_space(1))) unsigned long long gu64;
__device__ __forceinline__ float4 f4_of(u32x4_t v) { return __builtin_bit_cast(float4, v); }

// DPP moves: the other pair of the same row (lane l^1), and the row two
// lanes up (DIR > 0: lane l-2) or down (DIR < 0: lane l+2) by two
// wavefront shifts of one lane (lanes shifted in at the edge read zero)
__device__ __forceinline__ float dpp_pair(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xf, 0xf, true));
}
template <int DIR>
__device__ __forceinline__ float dpp_row(float v) {
  constexpr int ctl = DIR > 0 ? 0x138 : 0x130;   // wave_shr:1 / wave_shl:1
  const int a = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctl, 0xf, 0xf, true);
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(a, ctl, 0xf, 0xf, true));
}

// One FP32 sweep.  The rows of an instance are cut into W blocks of
// LU32_R; block w is one workgroup, lane l = 2r + h holding row
// ks + LU32_R*w + r and walking along j one column per step, row r one step
// behind row r-1 (forward; backward row r one step behind row r+1).
// Inside the workgroup, wave 0 computes; waves 1..NP stream the records of
// step t into LDS slot t % S, each keeping the loads of PD of its steps in
// flight; wave NP+1 reads the neighbour block's mail and puts the fill
// values into the same slot.  A slot holds four record rows and a fifth
// row of (u pair, fill pair) per lane: the producer writes the u half, the
// mail wave the fill half (meaningful in the edge row's two lanes, forward
// 0-1, backward 62-63: the neighbour block's edge row for the wavefront
// shift).  Each half has its own tag word.
// The record producers issue only cached buffer loads; the mail's
// system-scope polls stay in the mail wave.  A block with no neighbour
// upstream has zero fills, set up once when the workgroup starts.
// Hand-off protocol (LDS, workgroup scope; the LDS performs a CU's
// operations in one order that keeps each wave's issue order):
//   tags of slot t % S = t + 1  stored (release) by the producer after the
//                               records and u pair of step t, and by the
//                               mail wave after the fill pair
//   consumed = c                stored by the computing wave after it has
//                               issued the reads of steps < c, so a writer
//                               may fill slot t % S when consumed >= t-S+1
// The computing wave reads step t+2's slot while it computes step t and
// checks both tags when it starts step t+2: reads issued after the tags'
// read see what was stored before them (a stale tag means the slot is read
// again).
// Across workgroups: forward, block w waits on block w-1's mail; backward,
// on block w+1's.  A workgroup takes its block from a ticket counter of the
// instance (one atomic add at its start, never reset): ticket n is block
// n mod W of the instance's sweep number n / W, in the order the workgroups
// actually started, so the block a workgroup waits on belongs to a
// workgroup that is already running, whatever order the XCDs dispatch in.
// Every launch takes exactly W tickets per instance (workgroups x < W take
// one before anything else), and the sweep number + 1 is the mail tag.
constexpr int LU32_NV = 5;   // 16-byte rows per slot and lane
template <int DIR, int S, int NP, int PD>
__global__ __launch_bounds__(64 * (2 + NP)) void k_lu_sweep32(const DevMesh* __restrict__ Ms) {
  constexpr int NV = LU32_NV, R = LU32_R;
  extern __shared__ __attribute__((aligned(16))) double lds_raw[];
  __shared__ int consumed, abort_flag, st0, wb_s;
  __shared__ unsigned tag_s;
  u32x4_t* ring = reinterpret_cast<u32x4_t*>(lds_raw);             // [S][NV][64]
  int2* tags = reinterpret_cast<int2*>(ring + (size_t)S * NV * 64);  // [S]: records, fill
  const DevMesh& M = Ms[blockIdx.y];
  const LUGeom g = lu_geom(M);
  D2* lu = reinterpret_cast<D2*>(M.lu);
  if (threadIdx.x == 0) {
    consumed = 0;
    abort_flag = 0;
    wb_s = 0;
    tag_s = 0;
    if ((int)blockIdx.x < g.W) {
      unsigned long long* tk = reinterpret_cast<unsigned long long*>(lu + g.off_tk) + (DIR > 0 ? 0 : 1);
      const unsigned long long n = __hip_atomic_fetch_add(tk, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      wb_s = (int)(n % (unsigned long long)g.W);
      tag_s = (unsigned)(n / (unsigned long long)g.W) + 1u;
    }
    st0 = *M.status;
  }
  __syncthreads();
  if (st0 || (int)blockIdx.x >= g.W) return;
  const int wb = wb_s;
  const unsigned tag = tag_s;
  const int w = DIR > 0 ? wb : g.W - 1 - wb;   // row block of this workgroup
  // a block with no neighbour upstream takes zero fills at every step: its
  // fill words are zeroed once (nothing else writes them) and its fill tags
  // read as ready from the start
  const bool has_in = DIR > 0 ? w > 0 : w < g.W - 1;
  if (threadIdx.x < S) tags[threadIdx.x] = int2{0, has_in ? 0 : INT_MAX};
  if (!has_in && threadIdx.x < 2 * S) {
    const int el = (DIR > 0 ? 0 : 62) + (threadIdx.x & 1);
    unsigned* f = reinterpret_cast<unsigned*>(ring + (size_t)(threadIdx.x >> 1) * NV * 64 + 4 * 64 + el) + 2;
    *reinterpret_cast<u32x2_t*>(f) = u32x2_t{0u, 0u};
  }
  __syncthreads();
  const int NT = g.TW;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  // step-ordered result arrays (16 bytes per lane: value, tag, value,
  // tag): forward H (row NT-1-t for step t), backward H2 (row t); the
  // backward producers read H
  const long long hbytes = (long long)g.W * g.TW * 1024;
  const __amdgpu_buffer_rsrc_t hnd = uniform_rsrc(lu + g.off_h, hbytes);
  const __amdgpu_buffer_rsrc_t hout = DIR > 0 ? hnd : uniform_rsrc(lu + g.off_h2, hbytes);
  gu64* const hmail = (gu64*)(lu + (DIR > 0 ? g.off_h : g.off_h2));
  const int base = w * NT;
  if (wave == NP + 1) {
    // ---------------------------------------------------------- mail wave
    // The neighbour block's edge row at this block's step tt: forward, row
    // R-1 of block w-1 at column j = js+tt (its step tt+R-1); backward, row
    // 0 of block w+1 at column j = je-tt (its step tt+R-1).  One poll reads
    // 16 consecutive entries, lane l word l%4 of entry l/4; the entries from
    // the first one not yet written on wait for the next poll.  Steps
    // tt >= ni have no neighbour value and get zeros; a block with no
    // neighbour upstream has its zero fills from the start.
    if (!has_in) return;
    const int wn = w - DIR;
    const int nmail = g.ni;   // steps that take a neighbour value
    // lane l of a poll writes, for entry l/4, the pair of words l%4 and
    // l%4+1 (lanes with l%4 = 0, 2) into the edge row's lane of that pair
    const int dst = (DIR > 0 ? 0 : 62) + ((l >> 1) & 1);
    const bool writer = (l & 1) == 0;
    int tt = 0;
    bool ok = true;
    unsigned long long t_wait = rt_now();
    while (ok && tt < NT) {
      unsigned v = 0u, vn = 0u;
      int n = 0;   // entries ready from tt on
      if (tt < nmail) {
        const int e = tt + (l >> 2);
        const int ec = e < g.ni ? e : g.ni - 1;
        // the neighbour's step ec+R-1, stored at row NT-1-(ec+R-1)
        // forward and ec+R-1 backward; its edge lanes 62-63 forward, 0-1
        // backward (128 words per row)
        const int srow = DIR > 0 ? NT - 1 - (ec + R - 1) : ec + R - 1;
        const unsigned long long x =
            __hip_atomic_load(hmail + ((long long)wn * NT + srow) * 128 + (DIR > 0 ? 124 : 0) + (l & 3),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const bool ready = e < g.ni && (unsigned)(x >> 32) == tag;
        // entry i is ready when lanes 4i..4i+3 all are: count the ready
        // prefix of 4-lane groups
        const unsigned long long b = __ballot(ready);
        // bit 4i: entry i ready; every other bit set, so the lowest clear
        // bit is at 4 x (the first entry not ready)
        const unsigned long long nr =
            ~((b & (b >> 1) & (b >> 2) & (b >> 3) & 0x1111111111111111ull) | 0xEEEEEEEEEEEEEEEEull);
        n = nr ? (int)(__builtin_ctzll(nr) / 4) : 16;
        v = (unsigned)x;
        vn = __builtin_bit_cast(unsigned, __builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xf, 0xf, true));   // word of lane l^1
        if (n == 0) {
          if (__hip_atomic_load(&abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
          if (rt_now() - t_wait > LU32_WAIT) {
            __hip_atomic_store(&abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ok = false;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        t_wait = rt_now();
      } else {
        n = NT - tt < 16 ? NT - tt : 16;   // zero fills
      }
      // entries tt..tt+n-1 are all neighbour values (tt < nmail) or all
      // zeros.  The slot of step te is free once consumed >= te-S+1; every
      // entry whose slot is free is written in one pass (the consumer never
      // passes an unwritten entry, so consumed <= tt and at most S entries,
      // in distinct slots, are written at once), then their tags are set.
      const bool zero = tt >= nmail;
      for (int done = 0; done < n && ok;) {
        const int t1 = tt + done;
        const int c = __hip_atomic_load(&consumed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        int m = c + S - t1;
        if (m <= 0) {
          ok = spin_until_ge<4>(&consumed, t1 - S + 1, &abort_flag);
          continue;
        }
        m = m < n - done ? m : n - done;
        // mail entries: lanes 4i (words 0, 1) and 4i+2 (words 2, 3) of
        // entry i; zero entries: lanes 2i and 2i+1
        const int i = zero ? (l >> 1) : (l >> 2);
        if (i >= done && i < done + m && (zero || writer)) {
          u32x4_t* d = ring + (size_t)((tt + i) % S) * NV * 64 + 4 * 64;
          const int dl = zero ? (DIR > 0 ? 0 : 62) + (l & 1) : dst;
          *reinterpret_cast<u32x2_t*>(reinterpret_cast<unsigned*>(d + dl) + 2) =
              zero ? u32x2_t{0u, 0u} : u32x2_t{v, vn};
        }
        if (l >= done && l < done + m)
          __hip_atomic_store(&tags[(tt + l) % S].y, tt + l + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        done += m;
      }
      tt += n;
    }
    return;
  }
  if (wave > 0) {
    // ---------------------------------------------------------- producers
    // records in step layout: coalesced rows per step (lane l's cell at
    // step tt is (js+tt-r, ks+R*w+r) forward, the cell of forward step
    // NT-1-tt backward).  Entries of cells outside the grid hold zero
    // records: such lanes compute exact zeros from finite inputs.  Steps past
    // the end are clamped to the last one; the loop body has no conditional
    // loads.
    const __amdgpu_buffer_rsrc_t recf = uniform_rsrc(lu + g.off_f32, 288ll * g.W * g.TA * 16);
    const __amdgpu_buffer_rsrc_t recb = uniform_rsrc(lu + g.off_b32, 256ll * g.W * g.TA * 16);
    const __amdgpu_buffer_rsrc_t cbr = uniform_rsrc(lu + g.off_cb, (long long)g.W * g.TW * 512);
    const int r = l >> 1;
    const bool rowin = R * w + r < g.nk;
    // a step's records: four rows and the u pair (backward: s_fwd and cB
    // pairs), values in registers
    auto fetch = [&](int tt, u32x4_t (&v)[4], u32x2_t& u, u32x2_t& cb) {
      const int tc = tt < NT ? tt : NT - 1;
      if constexpr (DIR > 0) {
        const int so = (w * g.TA + tc) * 288 * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = __builtin_amdgcn_raw_buffer_load_b128(recf, l * 16, so + q * 1024, 0);
        u = __builtin_amdgcn_raw_buffer_load_b64(recf, l * 8, so + 4096, 0);
        cb = u32x2_t{0u, 0u};
      } else {
        const int so = (w * g.TA + NT - 1 - tc) * 256 * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = __builtin_amdgcn_raw_buffer_load_b128(recb, l * 16, so + q * 1024, 0);
        // the forward result pair (words 0 and 2 of the lane's 16 bytes)
        const u32x4_t h = __builtin_amdgcn_raw_buffer_load_b128(hnd, l * 16, (base + tc) * 1024, 0);
        u = u32x2_t{h.x, h.z};
        // cB of the cell, kept at its forward step NT-1-tc; zero outside
        // the grid (those entries are never written)
        const int jr = (g.ni - 1) - tc + (R - 1 - r);
        const bool cin = rowin && (unsigned)jr < (unsigned)g.ni;
        const u32x2_t c = __builtin_amdgcn_raw_buffer_load_b64(cbr, l * 8, (base + NT - 1 - tc) * 512, 0);
        cb = cin ? c : u32x2_t{0u, 0u};
      }
    };
    auto deliver = [&](int tt, const u32x4_t (&v)[4], u32x2_t u, u32x2_t cb) -> bool {
      if (tt >= NT) return true;
      // whole-vector casts: clang lowers __builtin_bit_cast of one element
      // of an ext_vector_type lvalue to a load from the vector's first
      // element, so the .y lane would read .x
      const F2 uu = __builtin_bit_cast(F2, u);
      float ux = uu.x, uy = uu.y;
      if constexpr (DIR < 0) {
        const F2 cc = __builtin_bit_cast(F2, cb);
        ux = ux + cc.x;
        uy = uy + cc.y;
      }
      if (!spin_until_ge<4>(&consumed, tt - S + 1, &abort_flag)) return false;
      u32x4_t* d = ring + (size_t)(tt % S) * NV * 64;
#pragma unroll
      for (int q = 0; q < 4; ++q) d[q * 64 + l] = v[q];
      *reinterpret_cast<u32x2_t*>(d + 4 * 64 + l) =
          u32x2_t{__builtin_bit_cast(unsigned, ux), __builtin_bit_cast(unsigned, uy)};
      if (l == 0) __hip_atomic_store(&tags[tt % S].x, tt + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      return true;
    };
    // producer wave p handles steps p-1, p-1+NP, ...; buffer k % PD holds the
    // k-th of them, fetched PD-1 deliveries ahead
    u32x4_t vb[PD][4];
    u32x2_t ub[PD], cbb[PD];
    const int t0 = wave - 1;
#pragma unroll
    for (int k = 0; k < PD - 1; ++k) fetch(t0 + k * NP, vb[k], ub[k], cbb[k]);
    bool ok = true;
    for (int tb = t0; tb < NT && ok; tb += PD * NP) {
#pragma unroll
      for (int k = 0; k < PD; ++k) {
        const int tt = tb + k * NP;
        const int kf = (k + PD - 1) % PD;
        fetch(tt + (PD - 1) * NP, vb[kf], ub[kf], cbb[kf]);
        if (ok) ok = deliver(tt, vb[k], ub[k], cbb[k]);
      }
    }
    return;
  }
  // ------------------------------------------------------------- compute
  // Lanes whose cell lies outside the grid compute exact zeros (their
  // records read as zeros), and a lane inside the grid takes its
  // neighbours' values only from cells inside the grid or through zero
  // matrices (the halo terms are folded into u and cB).
  __builtin_amdgcn_s_setprio(3);
  // the edge row's results are the neighbouring block's mail (entries
  // before step R-1 are never read); the opposite edge row takes the fill
  const bool fill_lane = DIR > 0 ? l < 2 : l >= 62;
  float a0 = 0.f, a1 = 0.f;   // this lane's pair of the previous step
  u32x4_t hv = {0u, tag, 0u, tag};   // the stored words; the tags stay put
  auto rd = [&](int slot, u32x4_t (&v)[NV], int& tv) {
    const u32x4_t* s = ring + (size_t)slot * NV * 64;
    const int2 tp = tags[slot];
    tv = tp.x < tp.y ? tp.x : tp.y;
    asm volatile("" ::: "memory");   // the tags' read before the records'
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = s[q * 64 + l];
  };
  auto spin_tags = [&](int slot, int target) -> bool {
    return spin_until_ge<1>(&tags[slot].x, target, &abort_flag) && spin_until_ge<1>(&tags[slot].y, target, &abort_flag);
  };
  bool ok = true;
  // step t from slot t % S (= slot, a constant once the loop below is
  // unrolled), its records read two steps earlier into C; reads step t+2
  // into N, the registers of step t-1
  auto step = [&](int slot, int t, u32x4_t (&C)[NV], int& tc, u32x4_t (&N)[NV], int& tn) {
    const int s2 = slot + 2 < S ? slot + 2 : slot + 2 - S;
    if (__builtin_expect(__builtin_amdgcn_readfirstlane(tc) < t + 1, 0)) {
      // the slot of step t was read before its writers finished: wait for
      // the tags and read it again
      ok = ok && spin_tags(slot, t + 1);
      rd(slot, C, tc);
    }
    // every third step (the last of a register-set rotation) frees the
    // slots up to step t.  Their reads were issued before this store (two
    // steps ago, or just above; the compiler barrier keeps them there) and
    // the LDS performs a CU's operations in one order that keeps each
    // wave's issue order, so they have read a slot before a writer that
    // sees this value refills it: no wait for reads still in flight.  The
    // writers then run up to S-3 steps ahead instead of S-1.
    if (slot % 3 == 2) {
      asm volatile("" ::: "memory");
      __hip_atomic_store(&consumed, t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    rd(s2, N, tn);   // past the last step: a harmless read of a stale slot
    const float4 x = f4_of(C[4]);
    // this row's other pair, and the neighbouring row's two pairs (the edge
    // row takes the fill)
    const float p0 = dpp_pair(a0), p1 = dpp_pair(a1);
    float k0 = dpp_row<DIR>(a0), k1 = dpp_row<DIR>(a1);
    k0 = fill_lane ? x.z : k0;
    k1 = fill_lane ? x.w : k1;
    const float kp0 = dpp_pair(k0), kp1 = dpp_pair(k1);
    const float4 A0 = f4_of(C[0]), A1 = f4_of(C[1]), B0 = f4_of(C[2]), B1 = f4_of(C[3]);
    float o0 = __builtin_fmaf(A0.x, a0, x.x), o1 = __builtin_fmaf(A1.x, a0, x.y);
    o0 = __builtin_fmaf(A0.y, a1, o0); o1 = __builtin_fmaf(A1.y, a1, o1);
    o0 = __builtin_fmaf(A0.z, p0, o0); o1 = __builtin_fmaf(A1.z, p0, o1);
    o0 = __builtin_fmaf(A0.w, p1, o0); o1 = __builtin_fmaf(A1.w, p1, o1);
    o0 = __builtin_fmaf(B0.x, k0, o0); o1 = __builtin_fmaf(B1.x, k0, o1);
    o0 = __builtin_fmaf(B0.y, k1, o0); o1 = __builtin_fmaf(B1.y, k1, o1);
    o0 = __builtin_fmaf(B0.z, kp0, o0); o1 = __builtin_fmaf(B1.z, kp0, o1);
    o0 = __builtin_fmaf(B0.w, kp1, o0); o1 = __builtin_fmaf(B1.w, kp1, o1);
    // the result of (step, lane), each value beside the tag: read by the
    // next kernels and, on the edge row, by the neighbour block as mail;
    // entries of cells outside the grid are never read as results.
    // Written through to memory (system scope).
    const int orow = DIR > 0 ? NT - 1 - t : t;
    hv.x = __builtin_bit_cast(unsigned, o0);
    hv.z = __builtin_bit_cast(unsigned, o1);
    __builtin_amdgcn_raw_buffer_store_b128(hv, hout, l * 16, (base + orow) * 1024, LU32_SYS);
    a0 = o0;
    a1 = o1;
  };
  u32x4_t C0[NV], C1[NV], C2[NV];
  int tc0 = 0, tc1 = 0, tc2 = 0;
  ok = spin_tags(0, 1);
  rd(0, C0, tc0);
  rd(1, C1, tc1);   // checked (and read again if early) at step 1
  // S steps per trip (S a multiple of 3), the three register sets rotating,
  // so that every slot index and LDS offset is a constant; the last,
  // partial trip checks each step against NT
  static_assert(S % 3 == 0, "the ring holds a multiple of 3 steps");
  int t0 = 0;
  for (; t0 + S <= NT && ok; t0 += S) {
#pragma unroll
    for (int i = 0; i < S; i += 3) {
      step(i, t0 + i, C0, tc0, C2, tc2);
      step(i + 1, t0 + i + 1, C1, tc1, C0, tc0);
      step(i + 2, t0 + i + 2, C2, tc2, C1, tc1);
    }
  }
  if (ok) {
#pragma unroll
    for (int i = 0; i < S; i += 3) {
      if (t0 + i < NT) step(i, t0 + i, C0, tc0, C2, tc2);
      if (t0 + i + 1 < NT) step(i + 1, t0 + i + 1, C1, tc1, C0, tc0);
      if (t0 + i + 2 < NT) step(i + 2, t0 + i + 2, C2, tc2, C1, tc1);
    }
  }
  if (!ok || __hip_atomic_load(&abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
    if (l == 0) atomicCAS(M.status, 0, 3);
}

__device__ __forceinline__ bool asq_bad(const DevConst& C, double q1, double q2, double q3, double q4) {
  double rho = q1;
  double qq = sq(q2) + sq(q3);
  double asq = C.ggm1 * (q4 - .5 * qq / rho) / rho;
  return asq <= 0.002 || asq + 1 == asq;
}

template <typename T>
// flags: 1 the q update and monitor follow here (Euler); 2 the SA model's
// laminar viscosity and velocities of every cell follow from the state
// written here (k_sa_prim's values)
__global__ __launch_bounds__(256) void k_lu_finish(const DevMesh* __restrict__ Ms, int flags) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int fuse_update = flags & 1, sa_prim = flags & 2;
  const int jd = M.jd, kd = M.kd, js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const LUGeom g = lu_geom(M);
  const LUArr B = lu_arr(M, g.off_b, g.nrec);
  const F4* H2 = reinterpret_cast<const F4*>(reinterpret_cast<const D2*>(M.lu) + g.off_h2);
  long long n2 = (long long)jd * kd;
  int bad = 0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    int j = 1 + (int)(i % jd), k = 1 + (int)(i / jd);
    const bool jin = j >= js && j <= je, kin = k >= ks && k <= ke;
    if constexpr (sizeof(T) == 4) {
      // halo ring: s*tscale (src/step.f:141-148), point Jacobi on the sides
      // (src/lhs.f:568-598); the FP32 prep folded these values unscaled-read
      if (!(jin && kin)) {
        CellLoad c = load_cell(M, i);
        double sc = c.ts;
        double d = (jin || kin) ? lu_of(M.c, c, 1. + 2. * M.c.epse).d : 1.0;
        for (int n = 0; n < 4; ++n) {
          double v = M.s[n][i] * sc;
          M.s[n][i] = (jin || kin) ? v * d : v;
        }
      }
    }
    double q1 = M.q[0][i], q2 = M.q[1][i], q3 = M.q[2][i], q4 = M.q[3][i];
    q_roundtrip(q1, q2, q3, q4, M.q[M.nq - 1][i]);
    if (jin && kin) {
      double dq[4];
      if constexpr (sizeof(T) == 8) {
        const long long e = lu_entry(M, g, j, k, -1);
        D2 a = *B.at(0, e), b = *B.at(1, e);
        dq[0] = a.x; dq[1] = a.y; dq[2] = b.x; dq[3] = b.y;
      } else {
        const long long e = lu_entry32(M, g, j, k, -1);
        const F4 a = H2[e], b = H2[e + 1];   // (value, tag) words
        dq[0] = (double)a.x; dq[1] = (double)a.z; dq[2] = (double)b.x; dq[3] = (double)b.z;
      }
      for (int n = 0; n < 4; ++n) M.s[n][i] = dq[n];
      if (fuse_update) {
        double ibp = (double)imax(M.iblank[i], 0);
        q1 = q1 + dq[0] * ibp; q2 = q2 + dq[1] * ibp; q3 = q3 + dq[2] * ibp; q4 = q4 + dq[3] * ibp;
      }
    }
    M.q[0][i] = q1; M.q[1][i] = q2; M.q[2][i] = q3; M.q[3][i] = q4;
    if (fuse_update) bad |= asq_bad(M.c, q1, q2, q3, q4);
    if (sa_prim) sa_prim_cell(M, i, q1, q2, q3, q4);
  }
  if (bad) atomicCAS(M.status, 0, 1);
}

// the q update (src/step.f:205-211) and monitor (src/step.f:332-392) in
// one pass: interior cells are updated, then every cell (halo included) is
// checked; each cell's update reads only that cell, so this is the update
// followed by the monitor
// take_turmu: the eddy viscosity the SA model left in turmu_n (it ran beside
// ilu2d) becomes turmu
__global__ __launch_bounds__(256) void k_update_monitor(const DevMesh* __restrict__ Ms, int take_turmu) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int jd = M.jd;
  long long n2 = (long long)jd * M.kd;
  int bad = 0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    const int j = 1 + (int)(i % jd), k = 1 + (int)(i / jd);
    double q1 = M.q[0][i], q2 = M.q[1][i], q3 = M.q[2][i], q4 = M.q[3][i];
    if (j >= M.js && j <= M.je && k >= M.ks && k <= M.ke) {
      double ibp = (double)imax(M.iblank[i], 0);
      q1 = q1 + M.s[0][i] * ibp; q2 = q2 + M.s[1][i] * ibp; q3 = q3 + M.s[2][i] * ibp; q4 = q4 + M.s[3][i] * ibp;
      M.q[0][i] = q1; M.q[1][i] = q2; M.q[2][i] = q3; M.q[3][i] = q4;
    }
    if (take_turmu) M.turmu[i] = M.turmu_n[i];
    bad |= asq_bad(M.c, q1, q2, q3, q4);
  }
  if (bad) atomicCAS(M.status, 0, 1);
}

// monitor (src/step.f:332-392) over all cells
__global__ __launch_bounds__(256) void k_monitor(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  long long n2 = (long long)M.jd * M.kd;
  int bad = 0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x)
    bad |= asq_bad(M.c, M.q[0][i], M.q[1][i], M.q[2][i], M.q[3][i]);
  if (bad) atomicCAS(M.status, 0, 1);
}

// ================================================================ geometry
// initia for one instance (src/init.f:2-177): find_coord_TS (setup.f:348),
// find_cellcenter (setup.f:111-345), metfv (setup.f:440-724),
// find_timemetrics_TS (setup.f:727-786), qzero (init.f:209-277), qdivj
// (init.f:180-206).  One block per instance, phases separated by barriers.
__global__ __launch_bounds__(256) void k_init_geometry(const DevMesh* __restrict__ Ms, int metrics_only) {
  const DevMesh& M = Ms[blockIdx.x];
  const DevConst& C = M.c;
  const int t = threadIdx.x, nt = blockDim.x;
  const int jd = M.jd, kd = M.kd, jmax = M.jmax, kmax = M.kmax;
  const long long nv2 = (long long)jmax * kmax, n2 = (long long)jd * kd;
  double* x = M.x;
  double* y = M.y;
  double* xv = M.xv;
  double* yv = M.yv;
  const double pi = 4.0 * atan(1.0);
  if (metrics_only) goto metrics;
  // find_coord_TS for instances 2..nspec (src/setup.f:348-393)
  if (M.nsp > 0 && C.timespectral) {
    double tspec = *M.tspec;
    if (C.motiontype == 1) {
      double theta = C.angmax * pi / 180. * (1. - cos(C.rotf * tspec));
      double ss = sin(theta), cs = cos(theta);
      for (long long w = t; w < nv2; w += nt) {
        double a = M.xv0[w], b = M.yv0[w];
        xv[w] = (a - C.xac) * cs + (b - C.yac) * ss + C.xac;
        yv[w] = (b - C.yac) * cs - (a - C.xac) * ss + C.yac;
      }
    } else if (C.motiontype == 2) {
      double cs = cos(C.rotf * tspec), ss = sin(C.rotf * tspec);
      for (long long w = t; w < nv2; w += nt) {
        double a = M.xv0[w], b = M.yv0[w];
        xv[w] = (a - C.xac) * cs - (b - C.yac) * ss + C.xac;
        yv[w] = (b - C.yac) * cs + (a - C.xac) * ss + C.yac;
      }
    }
    __syncthreads();
  }
  // cell centres
  for (long long w = t; w < n2; w += nt) { x[w] = 0.; y[w] = 0.; }
  __syncthreads();
#define XV(j, k) xv[vert(M, j, k)]
#define YV(j, k) yv[vert(M, j, k)]
#define XC(j, k) x[cell(M, j, k)]
#define YC(j, k) y[cell(M, j, k)]
  {
    int nj = M.je - M.js + 1, nk = M.ke - M.ks + 1;
    for (long long w = t; w < (long long)nj * nk; w += nt) {
      int j = M.js + (int)(w % nj), k = M.ks + (int)(w / nj);
      XC(j, k) = 0.25 * (XV(j, k) + XV(j - 1, k) + XV(j, k - 1) + XV(j - 1, k - 1));
      YC(j, k) = 0.25 * (YV(j, k) + YV(j - 1, k) + YV(j, k - 1) + YV(j - 1, k - 1));
    }
  }
  __syncthreads();
  for (int k = M.ks + t; k <= M.ke; k += nt) {
    int j = M.js - 1;
    XC(j, k) = XC(j + 1, k) - 0.5 * (XV(j + 1, k) - XV(j, k) + XV(j + 1, k - 1) - XV(j, k - 1));
    YC(j, k) = YC(j + 1, k) - 0.5 * (YV(j + 1, k) - YV(j, k) + YV(j + 1, k - 1) - YV(j, k - 1));
    j = M.je + 1;
    XC(j, k) = XC(j - 1, k) - 0.5 * (XV(j - 2, k) - XV(j - 1, k) + XV(j - 2, k - 1) - XV(j - 1, k - 1));
    YC(j, k) = YC(j - 1, k) - 0.5 * (YV(j - 2, k) - YV(j - 1, k) + YV(j - 2, k - 1) - YV(j - 1, k - 1));
  }
  for (int j = M.js + t; j <= M.je; j += nt) {
    int k = M.ks - 1;
    XC(j, k) = XC(j, k + 1) - 0.5 * (XV(j, k + 1) - XV(j, k) + XV(j - 1, k + 1) - XV(j - 1, k));
    YC(j, k) = YC(j, k + 1) - 0.5 * (YV(j, k + 1) - YV(j, k) + YV(j - 1, k + 1) - YV(j - 1, k));
    k = M.ke + 1;
    XC(j, k) = XC(j, k - 1) - 0.5 * (XV(j, k - 2) - XV(j, k - 1) + XV(j - 1, k - 2) - XV(j - 1, k - 1));
    YC(j, k) = YC(j, k - 1) - 0.5 * (YV(j, k - 2) - YV(j, k - 1) + YV(j - 1, k - 2) - YV(j - 1, k - 1));
  }
  __syncthreads();
  if (t == 0) {
    int j = M.js - 1, k = M.ks - 1;
    XC(j, k) = 2 * XC(j + 1, k) - XC(j + 2, k);
    YC(j, k) = 2 * YC(j + 1, k) - YC(j + 2, k);
    k = M.ke + 1;
    XC(j, k) = 2 * XC(j + 1, k) - XC(j + 2, k);
    YC(j, k) = 2 * YC(j + 1, k) - YC(j + 2, k);
    j = M.je + 1; k = M.ks - 1;
    XC(j, k) = 2 * XC(j, k + 1) - XC(j, k + 2);
    YC(j, k) = 2 * YC(j, k + 1) - YC(j, k + 2);
    k = M.ke + 1;
    XC(j, k) = 2 * XC(j, k - 1) - XC(j, k - 2);
    YC(j, k) = 2 * YC(j, k - 1) - YC(j, k - 2);
  }
  __syncthreads();
  for (int ib = 0; ib < M.nbc; ++ib) {
    const DevBC& b = M.bc[ib];
    int js = b.js, je = b.je, ks = b.ks, ke = b.ke, idir = b.idir;
    int nj = je - js + 1, nk = ke - ks + 1;
    if (b.type == 22) {
      for (int w = t; w < nj * nk; w += nt) {
        int j = js + w % nj, k = ks + w / nj;
        if (idir == 1 || idir == -1) {
          int jc = periodic_src(idir, js, je, ks, ke, jd, kd, j);
          XC(j, k) = XC(jc, k); YC(j, k) = YC(jc, k);
        } else {
          int kc = periodic_src(idir, js, je, ks, ke, jd, kd, k);
          XC(j, k) = XC(j, kc); YC(j, k) = YC(j, kc);
        }
      }
    } else if (b.type == 51 && (idir == 2 || idir == -2)) {
      int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
      for (int j = js + t; j <= je; j += nt) {
        int jj = jd - j + 1;
        double a = XC(jj, k1), bb = YC(jj, k1), c = XC(j, k1), d = YC(j, k1);
        XC(j, k) = a; YC(j, k) = bb; XC(jj, k) = c; YC(jj, k) = d;
      }
    }
    __syncthreads();
  }
  // ---- metfv
metrics:
  double* qj = M.q[M.nq - 1];
  for (long long w = t; w < (long long)(jmax - 1) * (kmax - 1); w += nt) {
    int j = 2 + (int)(w % (jmax - 1)), k = 2 + (int)(w / (jmax - 1));
    long long i = cell(M, j, k);  // jc = j, kc = k with one halo
    double dx1 = XV(j, k) - XV(j - 1, k - 1);
    double dy1 = YV(j, k) - YV(j - 1, k - 1);
    double dx2 = XV(j - 1, k) - XV(j, k - 1);
    double dy2 = YV(j - 1, k) - YV(j, k - 1);
    double a3 = 0.5 * (dx1 * dy2 - dx2 * dy1);
    double J = 1. / a3;
    qj[i] = J;
    double xx = 0.5 * (YV(j, k) - YV(j, k - 1) + YV(j - 1, k) - YV(j - 1, k - 1));
    double xy = -0.5 * (XV(j, k) - XV(j, k - 1) + XV(j - 1, k) - XV(j - 1, k - 1));
    double yx = -0.5 * (YV(j, k) - YV(j - 1, k) + YV(j, k - 1) - YV(j - 1, k - 1));
    double yy = 0.5 * (XV(j, k) - XV(j - 1, k) + XV(j, k - 1) - XV(j - 1, k - 1));
    M.xx[i] = J * xx; M.xy[i] = J * xy; M.yx[i] = J * yx; M.yy[i] = J * yy;
  }
  __syncthreads();
#define CPY(dst, src) { qj[dst] = qj[src]; M.xx[dst] = M.xx[src]; M.xy[dst] = M.xy[src]; \
    M.yx[dst] = M.yx[src]; M.yy[dst] = M.yy[src]; }
  for (int k = M.ks + t; k <= M.ke; k += nt) {
    CPY(cell(M, M.js - 1, k), cell(M, M.js, k));
    CPY(cell(M, M.je + 1, k), cell(M, M.je, k));
  }
  __syncthreads();
  for (int j = M.js - 1 + t; j <= M.je + 1; j += nt) {
    CPY(cell(M, j, M.ks - 1), cell(M, j, M.ks));
    CPY(cell(M, j, M.ke + 1), cell(M, j, M.ke));
  }
  __syncthreads();
  for (int ib = 0; ib < M.nbc; ++ib) {
    const DevBC& b = M.bc[ib];
    int js = b.js, je = b.je, ks = b.ks, ke = b.ke, idir = b.idir;
    int nj = je - js + 1, nk = ke - ks + 1;
    if (b.type == 22) {
      for (int w = t; w < nj * nk; w += nt) {
        int j = js + w % nj, k = ks + w / nj;
        if (idir == 1 || idir == -1) {
          int jc = periodic_src(idir, js, je, ks, ke, jd, kd, j);
          CPY(cell(M, j, k), cell(M, jc, k));
        } else {
          int kc = periodic_src(idir, js, je, ks, ke, jd, kd, k);
          CPY(cell(M, j, k), cell(M, j, kc));
        }
      }
    } else if (b.type == 51 && (idir == 2 || idir == -2)) {
      int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
      for (int j = js + t; j <= je; j += nt) {
        int jj = jd - j + 1;
        long long a = cell(M, j, k), bb = cell(M, jj, k1), cc = cell(M, jj, k), d = cell(M, j, k1);
        double v0 = qj[bb], v1 = -M.xx[bb], v2 = -M.xy[bb], v3 = -M.yx[bb], v4 = -M.yy[bb];
        double w0 = qj[d], w1 = -M.xx[d], w2 = -M.xy[d], w3 = -M.yx[d], w4 = -M.yy[d];
        qj[a] = v0; M.xx[a] = v1; M.xy[a] = v2; M.yx[a] = v3; M.yy[a] = v4;
        qj[cc] = w0; M.xx[cc] = w1; M.xy[cc] = w2; M.yx[cc] = w3; M.yy[cc] = w4;
      }
    }
    __syncthreads();
  }
#undef CPY
  if (metrics_only) return;
  // ---- find_timemetrics_TS (src/setup.f:727-786)
  if (C.timespectral) {
    double tspec = *M.tspec, rotf = C.rotf, am = C.angmax;
    if (C.motiontype == 1) {
      for (long long i = t; i < n2; i += nt) {
        M.ug[i] = am * pi / 180. * rotf * (y[i] - C.yac) * sin(rotf * tspec);
        M.vg[i] = -am * pi / 180. * rotf * (x[i] - C.xac) * sin(rotf * tspec);
      }
      for (long long i = t; i < nv2; i += nt) {
        M.ugv[i] = am * pi / 180. * rotf * (yv[i] - C.yac) * sin(rotf * tspec);
        M.vgv[i] = -am * pi / 180. * rotf * (xv[i] - C.xac) * sin(rotf * tspec);
      }
    } else if (C.motiontype == 2) {
      for (long long i = t; i < n2; i += nt) { M.ug[i] = -rotf * (y[i] - C.yac); M.vg[i] = rotf * (x[i] - C.xac); }
      for (long long i = t; i < nv2; i += nt) { M.ugv[i] = -rotf * (yv[i] - C.yac); M.vgv[i] = rotf * (xv[i] - C.xac); }
    }
  } else {
    for (long long i = t; i < n2; i += nt) { M.ug[i] = 0.; M.vg[i] = 0.; }
  }
  // ---- qzero + qdivj
  double ruinf = C.rinf * C.uinf, rvinf = C.rinf * C.vinf;
  for (long long i = t; i < n2; i += nt) {
    double J = qj[i];
    M.q[0][i] = C.rinf / J;
    M.q[1][i] = ruinf / J;
    M.q[2][i] = rvinf / J;
    M.q[3][i] = C.einf / J;
    if (C.iturb == 1) M.q[4][i] = C.vnuinf;
    if (C.iturb == 2) {  // src/init.f:246-254
      M.q[4][i] = C.rinf * C.tkeinf;
      M.q[5][i] = C.rinf * C.tomegainf;
    }
    M.turmu[i] = 0.;
    M.iblank[i] = 1;
    M.bt[i] = 1.;
    M.tscale[i] = 0.;
    for (int n = 0; n < M.nv; ++n) M.s[n][i] = 0.;
  }
#undef XV
#undef YV
#undef XC
#undef YC
}

// ============================================================ pitching grid
// pitch (src/motion.f:546-632) at time step istep of the run: the cell
// centres and vertices turn about (xac, yac) by this step's angle, the grid
// velocity of every cell follows from the turn (first order, or second order
// with the previous turn when ntac > 1), and the vertex levels shift (xole
// <- xold <- xv).  The metrics follow from the turned vertices (metfv).
__global__ __launch_bounds__(256) void k_pitch(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const DevConst& C = M.c;
  const double pi = 4.0 * atan(1.0);
  const double dt = C.h;
  const int istep = *M.istep;
  const double dtheta = -C.angmax * (cos(C.rotf * istep * dt) - cos(C.rotf * (istep - 1) * dt));
  const double sn = sin(dtheta * pi / 180.), cn = cos(dtheta * pi / 180.);
  const double snold = sin(-1. * dtheta * pi / 180.), cnold = cos(-1. * dtheta * pi / 180.);
  const long long n2 = (long long)M.jd * M.kd, nv2 = (long long)M.jmax * M.kmax;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    const double x = M.x[i], y = M.y[i];
    const double xnew = (x - C.xac) * cn + (y - C.yac) * sn + C.xac;
    const double ynew = -(x - C.xac) * sn + (y - C.yac) * cn + C.yac;
    if (C.ntac == 1) {
      M.ug[i] = (xnew - x) / dt;
      M.vg[i] = (ynew - y) / dt;
    } else {
      const double xo = (x - C.xac) * cnold + (y - C.yac) * snold + C.xac;
      const double yo = -(x - C.xac) * snold + (y - C.yac) * cnold + C.yac;
      M.ug[i] = (1.5 * xnew - 2. * x + 0.5 * xo) / dt;
      M.vg[i] = (1.5 * ynew - 2. * y + 0.5 * yo) / dt;
    }
    M.x[i] = xnew;
    M.y[i] = ynew;
  }
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < nv2; i += 256ll * gridDim.x) {
    M.xole[i] = M.xold[i];
    M.yole[i] = M.yold[i];
    const double xo = M.xv[i], yo = M.yv[i];
    M.xold[i] = xo;
    M.yold[i] = yo;
    M.xv[i] = (xo - C.xac) * cn + (yo - C.yac) * sn + C.xac;
    M.yv[i] = -(xo - C.xac) * sn + (yo - C.yac) * cn + C.yac;
  }
}

// ================================================================== forces
// compute_forces/force2d (src/forces.f:2-398) for one instance: a single
// block reduces the wall integrals in a fixed order.
__global__ __launch_bounds__(256) void k_forces(const DevMesh* __restrict__ Ms, int mi, double* out) {
  const DevMesh& M = Ms[mi];
  const DevConst& C = M.c;
  __shared__ double red[8][256];
  const int t = threadIdx.x;
  const int k = M.ks, k2 = k + 1, k3 = k2 + 1;
  const int j1 = M.jtail1, j2 = M.jtail2;
  const double* q1 = M.q[0];
  const double* q2 = M.q[1];
  const double* q3 = M.q[2];
  const double* q4 = M.q[3];
  const double* qj = M.q[M.nq - 1];
  const double fmm = C.fmm, halfv = C.half;
  double gm1 = C.gm1, gamma = C.gamma;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // cn, cc, cmle, cpinv, cnv, ccv, cmlev, cpv
  double amu = C.rinf * 1. / C.rey;
  double uinf2 = sq(fmm);
  auto U = [&](int j, int kk) { long long i = cell(M, j, kk); return q2[i] / q1[i]; };
  auto V = [&](int j, int kk) { long long i = cell(M, j, kk); return q3[i] / q1[i]; };
  for (int j = j1 + t; j <= j2; j += 256) {
    long long i = cell(M, j, k), im1 = cell(M, j, k - 1);
    double rr = 1. / q1[i];
    double rho = q1[i] * qj[i];
    double u = q2[i] * rr, v = q3[i] * rr, e = q4[i] * qj[i];
    double vsq = u * u + v * v;
    double pp = gm1 * (e - rho * vsq / 2.);
    double pp1 = pp / C.pinf;
    double cp = 2. * (pp1 - 1.) / (gamma * sq(fmm));
    double zz = cp;
    rr = 1. / q1[im1];
    rho = q1[im1] * qj[im1];
    u = q2[im1] * rr; v = q3[im1] * rr; e = q4[im1] * qj[im1];
    vsq = u * u + v * v;
    pp = gm1 * (e - rho * vsq / 2.);
    pp1 = pp / C.pinf;
    cp = 2. * (pp1 - 1.) / (gamma * sq(fmm));
    zz = 0.5 * (zz + cp);
    int jv = j - 1, kv = k - 1;
    double xa = M.xv[vert(M, jv, kv)], xb = M.xv[vert(M, jv + 1, kv)];
    double ya = M.yv[vert(M, jv, kv)], yb = M.yv[vert(M, jv + 1, kv)];
    acc[0] += -zz * (xb - xa);
    acc[1] += zz * (yb - ya);
    acc[2] += zz * (0.5 * (xb + xa) * (xb - xa) + 0.5 * (yb + ya) * (yb - ya));
    double uu = q2[i] / q1[i], vv = q3[i] / q1[i];
    double um = q2[im1] / q1[im1], vm = q3[im1] / q1[im1];
    acc[3] += 0.5 * (uu + um) * zz * (yb - ya) - 0.5 * (vv + vm) * zz * (xb - xa);
    if (!C.invisc) {
      double uxi, vxi;
      if (j == j1) { uxi = U(j + 1, k) - U(j, k); vxi = V(j + 1, k) - V(j, k); }
      else if (j == j2) { uxi = U(j, k) - U(j - 1, k); vxi = V(j, k) - V(j - 1, k); }
      else { uxi = .5 * (U(j + 1, k) - U(j - 1, k)); vxi = .5 * (V(j + 1, k) - V(j - 1, k)); }
      // -1.5*q(j,k,2)/q(j,k,1) is (-1.5*q2)/q1 (src/forces.f:299-302)
      long long i1 = cell(M, j, k), i2 = cell(M, j, k2), i3 = cell(M, j, k3);
      double ueta = -1.5 * q2[i1] / q1[i1] + 2. * q2[i2] / q1[i2] - .5 * q2[i3] / q1[i3];
      double veta = -1.5 * q3[i1] / q1[i1] + 2. * q3[i2] / q1[i2] - .5 * q3[i3] / q1[i3];
      double xix = M.xx[i], xiy = M.xy[i], etax = M.yx[i], etay = M.yy[i];
      double sfdiv = 2. / 3 * (uxi * xix + ueta * etax + vxi * xiy + veta * etay);
      double sxx = amu * (2. * (uxi * xix + ueta * etax) - sfdiv);
      double syy = amu * (2. * (vxi * xiy + veta * etay) - sfdiv);
      double tauw = amu * ((uxi * xiy + ueta * etay) + (vxi * xix + veta * etax));
      sxx = sxx / (.5 * C.rinf * uinf2);
      syy = syy / (.5 * C.rinf * uinf2);
      double zv = tauw / (.5 * C.rinf * uinf2);
      acc[5] += zv * (xb - xa) - sxx * (yb - ya);
      acc[4] += -zv * (yb - ya) + syy * (xb - xa);
      acc[6] += zv * (0.5 * (xb + xa) * (yb - ya) + 0.5 * (yb + ya) * (xb - xa))
                - 0.5 * (xb + xa) * (xb - xa) * syy - 0.5 * (yb + ya) * (yb - ya) * sxx;
      acc[7] += 0.5 * (uu + um) * zv * (xb - xa) - 0.5 * (uu + um) * sxx * (yb - ya)
                - 0.5 * (vv + vm) * zv * (yb - ya) + 0.5 * (vv + vm) * syy * (xb - xa);
    }
  }
  for (int n = 0; n < 8; ++n) red[n][t] = acc[n];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
      for (int n = 0; n < 8; ++n) red[n][t] += red[n][t + off];
    __syncthreads();
  }
  if (t == 0) {
    double cn = red[0][0], cc = red[1][0], cmle = red[2][0], cpinv = red[3][0];
    double cnv = red[4][0], ccv = red[5][0], cmlev = red[6][0], cpv = red[7][0];
    cn = cn - halfv * cn; cmle = cmle - halfv * cmle; cc = cc + halfv * cc;
    cnv = cnv - halfv * cnv; cmlev = cmlev - halfv * cmlev; ccv = ccv + halfv * ccv;
    double cfx = cc + ccv, cfy = cn + cnv, cm0 = cmle + cmlev;
    double cpower = (cpinv + cpv) / fmm;
    double ca = C.ca, sa = C.sa;
    out[0] = -cfx * sa + cfy * ca;
    out[1] = cfx * ca + cfy * sa;
    out[2] = cm0 - 0.0 * cfx + 0.25 * cfy;
    out[3] = cpower;
    out[4] = cfx;
    out[5] = cfy;
  }
}

// ================================================================ launchers
static inline int nblk(long long n, int per) {
  long long b = (n + per - 1) / per;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

hipError_t launch_init_geometry(const DevMesh* dM, int nmi, hipStream_t st, int metrics_only) {
  hipLaunchKernelGGL(k_init_geometry, dim3(nmi), dim3(256), 0, st, dM, metrics_only);
  return hipGetLastError();
}

hipError_t launch_pitch(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_pitch, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM);
  return hipGetLastError();
}

hipError_t launch_time_step(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st, int* istep0,
                            const int* status, int* stepc, int ncase, int itn) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_time_step, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM, istep0, status, stepc, ncase,
                     itn);
  return hipGetLastError();
}

hipError_t launch_open_step(const int* status, int* istep0, int* stepc, int* istep, int ncase, hipStream_t st) {
  hipLaunchKernelGGL(k_open_step, dim3((ncase + 255) / 256), dim3(256), 0, st, status, istep0, stepc, istep, ncase);
  return hipGetLastError();
}

hipError_t launch_stqol(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_stqol, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM);
  return hipGetLastError();
}

hipError_t launch_newton(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_newton, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM);
  return hipGetLastError();
}

hipError_t launch_copy_doubles(double* dst, const double* src, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_doubles, dim3((n + 255) / 256), dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_bc(const DevMesh* dM, const DevMesh* hM, int nmi, int zero_s, hipStream_t st) {
  int nbc = 1, nph = 1;
  for (int i = 0; i < nmi; ++i) {
    nbc = hM[i].nbc > nbc ? hM[i].nbc : nbc;
    nph = hM[i].nbc_phase > nph ? hM[i].nbc_phase : nph;
  }
  for (int ph = 0; ph < nph; ++ph) {
    hipLaunchKernelGGL(k_bc, dim3(nbc, nmi), dim3(BC_THREADS), 0, st, dM, zero_s, ph);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_residual(const DevMesh* dM, const ResTile* tiles, int ntiles, int do_norms, int weno,
                           hipStream_t st) {
  if (weno)
    hipLaunchKernelGGL(k_residual<3>, dim3(ntiles), dim3(RES_THREADS), 0, st, dM, tiles, do_norms);
  else
    hipLaunchKernelGGL(k_residual<2>, dim3(ntiles), dim3(RES_THREADS), 0, st, dM, tiles, do_norms);
  return hipGetLastError();
}

hipError_t launch_norms(const DevMesh* dM, const ResTile* tiles, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL(k_norms, dim3(ntiles), dim3(RES_THREADS), 0, st, dM, tiles);
  return hipGetLastError();
}

// convergence of every case: status 4 once the rsum of all its instances is
// at or below its tolerance (tol <= 0: never).  inst[2c], inst[2c+1]: first
// and one-past-last instance record of case c.
__global__ __launch_bounds__(256) void k_conv_check(const DevMesh* __restrict__ Ms, const int* __restrict__ inst,
                                                    const double* __restrict__ tol, int* status, int ncase) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncase || status[c] != 0 || !(tol[c] > 0.)) return;
  double mx = 0.;
  bool finite = true;
  for (int i = inst[2 * c]; i < inst[2 * c + 1]; ++i) {
    double r = Ms[i].norms[0];
    finite = finite && r == r && r - r == 0.;
    mx = r > mx ? r : mx;
  }
  // a residual that is not a finite number stops the case as the norm check
  // does a blown-up one (status 2), never as converged
  if (!finite) status[c] = 2;
  else if (mx <= tol[c]) status[c] = 4;
}

hipError_t launch_conv_check(const DevMesh* dM, const int* inst, const double* tol, int* status, int ncase,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_conv_check, dim3((ncase + 255) / 256), dim3(256), 0, st, dM, inst, tol, status, ncase);
  return hipGetLastError();
}

hipError_t launch_norms_final(const DevMesh* dM, int nmi, hipStream_t st, int rbc) {
  hipLaunchKernelGGL(k_norms_final, dim3(nmi), dim3(RES_THREADS), 0, st, dM, rbc);
  return hipGetLastError();
}

hipError_t launch_tsrhs(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_tsrhs, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM);
  return hipGetLastError();
}

hipError_t launch_rbc(const DevMesh* dM, int nmi, hipStream_t st) {
  hipLaunchKernelGGL(k_rbc, dim3(nmi), dim3(BC_THREADS), 0, st, dM);
  return hipGetLastError();
}

template <typename T, int RL, int D>
static hipError_t launch_lu_t(const DevMesh* dM, const DevMesh* hM, int nmi, int fuse_update, Marks* mk,
                              hipStream_t st) {
  long long mx = 0;
  int nimax = 0;
  for (int i = 0; i < nmi; ++i) {
    mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
    nimax = hM[i].je - hM[i].js + 1 > nimax ? hM[i].je - hM[i].js + 1 : nimax;
  }
  size_t lds = lu_lds_bytes<T>(nimax);
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_lu_sweep<T, 1, RL, D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void*)k_lu_sweep<T, -1, RL, D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_lu_prep, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_PREP, st);
  hipLaunchKernelGGL((k_lu_sweep<T, 1, RL, D>), dim3(nmi), dim3(64), lds, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_FORWARD, st);
  hipLaunchKernelGGL((k_lu_sweep<T, -1, RL, D>), dim3(nmi), dim3(64), lds, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_BACKWARD, st);
  if (fuse_update & 4) return hipSuccess;
  hipLaunchKernelGGL(k_lu_finish<T>, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM, fuse_update);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_FINISH, st);
  return hipSuccess;
}

// LDS ring slots of the FP32 sweeps (LU32_NV KiB each and two tag words): 14
// slots, so two workgroups fit one CU's LDS; producer waves and the steps
// each keeps in flight
// the FP32 sweep's instantiations (ring steps S, producer waves NP,
// deliveries in flight per producer PD): the same arithmetic in the same
// order, so the choice (LaunchCfg::sweep, the autotuner's) changes no result
constexpr size_t lu32_lds(int S) { return (size_t)S * LU32_NV * 1024 + 16 * ((S * 8 + 15) / 16); }

template <int S, int NP, int PD>
static hipError_t launch_sweeps32(const DevMesh* dM, int wmax, int nmi, Marks* mk, hipStream_t st) {
  constexpr size_t lds = lu32_lds(S);
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_lu_sweep32<1, S, NP, PD>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    (void)hipFuncSetAttribute((const void*)k_lu_sweep32<-1, S, NP, PD>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    attr_set = true;
  }
  hipLaunchKernelGGL((k_lu_sweep32<1, S, NP, PD>), dim3(wmax, nmi), dim3(64 * (2 + NP)), lds, st, dM);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_FORWARD, st);
  hipLaunchKernelGGL((k_lu_sweep32<-1, S, NP, PD>), dim3(wmax, nmi), dim3(64 * (2 + NP)), lds, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_BACKWARD, st);
  return hipSuccess;
}

static hipError_t launch_lu32(const DevMesh* dM, const DevMesh* hM, int nmi, int fuse_update, Marks* mk,
                              hipStream_t st, int sweep) {
  long long mx = 0, nc = 0;
  int wmax = 1;
  for (int i = 0; i < nmi; ++i) {
    mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
    int nk = hM[i].ke - hM[i].ks + 1, ni = hM[i].je - hM[i].js + 1;
    const LUGeom gg = lu_geom(ni, nk, hM[i].lu_rl);
    const long long nprep = (long long)gg.W * (LU32_R / 8) * ((gg.TA + LP_STEPS - 1) / LP_STEPS);   // prep tiles
    nc = nprep > nc ? nprep : nc;
    wmax = gg.W > wmax ? gg.W : wmax;
  }
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_lu_prep32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LP_LDS);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_lu_prep32, dim3((unsigned)nc, nmi), dim3(64 * LP_WAVES), LP_LDS, st, dM);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_PREP, st);
  switch (sweep) {
    case 0: e = launch_sweeps32<12, 3, 5>(dM, wmax, nmi, mk, st); break;
    case 1: e = launch_sweeps32<18, 3, 5>(dM, wmax, nmi, mk, st); break;
    case 2: e = launch_sweeps32<12, 4, 6>(dM, wmax, nmi, mk, st); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  if (fuse_update & 4) return hipSuccess;
  hipLaunchKernelGGL(k_lu_finish<float>, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM, fuse_update);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_LU_FINISH, st);
  return hipSuccess;
}

hipError_t launch_lu_finish(const DevMesh* dM, const DevMesh* hM, int nmi, int flags, int fp32, hipStream_t st) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  if (fp32) hipLaunchKernelGGL(k_lu_finish<float>, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM, flags);
  else hipLaunchKernelGGL(k_lu_finish<double>, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM, flags);
  return hipGetLastError();
}

// rows per lane rl = ceil(rows / 64) of the case (all instances share it)
hipError_t launch_lu(const DevMesh* dM, const DevMesh* hM, int nmi, int fuse_update, int fp32, Marks* mk,
                     hipStream_t st, int sweep) {
  const int rl = hM[0].lu_rl;
  if (fp32) {
    return launch_lu32(dM, hM, nmi, fuse_update, mk, st, sweep);
  } else {
    // prefetch depth D chosen so the record ring fits the register file
    switch (rl) {
      case 1: return launch_lu_t<double, 1, 3>(dM, hM, nmi, fuse_update, mk, st);
      case 2: return launch_lu_t<double, 2, 2>(dM, hM, nmi, fuse_update, mk, st);
      case 3: return launch_lu_t<double, 3, 1>(dM, hM, nmi, fuse_update, mk, st);
      case 4: return launch_lu_t<double, 4, 1>(dM, hM, nmi, fuse_update, mk, st);
      // tall meshes (Cartesian backgrounds of overset cases); the FP32 sweep
      // has no such bound
      case 5: return launch_lu_t<double, 5, 1>(dM, hM, nmi, fuse_update, mk, st);
      case 6: return launch_lu_t<double, 6, 1>(dM, hM, nmi, fuse_update, mk, st);
      case 7: return launch_lu_t<double, 7, 1>(dM, hM, nmi, fuse_update, mk, st);
      case 8: return launch_lu_t<double, 8, 1>(dM, hM, nmi, fuse_update, mk, st);
    }
  }
  return hipErrorInvalidValue;
}

size_t lu_lds_need(int ni, int fp32) {
  return fp32 ? lu_lds_bytes<float>(ni) : lu_lds_bytes<double>(ni);
}

size_t lu_workspace_bytes(int jd, int kd, int rl) {
  return 16 * (size_t)lu_geom(jd - 2, kd - 2, rl).total;
}

hipError_t launch_update_monitor(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st, int take_turmu) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_update_monitor, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM, take_turmu);
  return hipGetLastError();
}
hipError_t launch_monitor(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st) {
  long long mx = 0;
  for (int i = 0; i < nmi; ++i) mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
  hipLaunchKernelGGL(k_monitor, dim3(nblk(mx, 256), nmi), dim3(256), 0, st, dM);
  return hipGetLastError();
}

hipError_t launch_forces(const DevMesh* dM, int mi, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_forces, dim3(1), dim3(256), 0, st, dM, mi, out);
  return hipGetLastError();
}

}  // namespace ot2d
