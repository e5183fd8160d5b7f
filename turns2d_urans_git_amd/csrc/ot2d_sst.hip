// gfx950 kernels of the k-omega SST model (komegasst, src/komegasst.f:2-187,
// itrans = 0).
//
//   k_sa_prim (shared with SA)   laminar viscosity, u, v (all cells)
//   k_sst_bc                     komegabc (src/komegasst.f:906-1434), segments in
//                                order, one block per instance; with eddy=1 the
//                                eddy viscosity of the halo ring follows
//   k_sst_blend                  kwvortic, turbstress, crossdiff, blend
//                                (src/komegasst.f:270-461) per interior cell and
//                                findturm (190-215) for every cell
//   k_sst_rhs                    komegarhslhs (464-903) in gather form: each
//                                interior cell assembles its rows of the four
//                                tridiagonal systems, scaled by the pseudo time step
//   k_sst_solve_j / _k           lsolvej2 / lsolvek2 (1435-1518), one wave per
//                                line and equation; the k solve applies the update
//                                with its clip (146-172)
//   k_sst_turm                   findturm after the update
//
// The reference reads F1, F2 and strain on the halo ring without writing
// them there (blend and kwvortic fill the interior only; findturm and
// komegadiffus read every cell).  Those values are taken as zero, which is
// what the reference computes when its allocations come back zero-filled.
// Operand order follows the reference expressions (-ffp-contract=off).
#include "ot2d_dev.h"

namespace ot2d {

namespace {
constexpr double KW_sigmak1 = 0.85, KW_sigmaw1 = 0.5, KW_beta1 = 0.075, KW_sigmak2 = 1.0, KW_sigmaw2 = 0.856,
                 KW_beta2 = 0.0828, KW_betastar = 0.09;
constexpr double KW_gamma1 = 5. / 9, KW_gamma2 = 0.44;

__device__ __forceinline__ double sqk(double a) { return a * a; }
__device__ __forceinline__ int imk(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ int kw_periodic_src(int idir, int js, int je, int ks, int ke, int jd, int kd, int jk) {
  if (idir == 1) { int jj = je - js + 1; return jd - 2 * jj + (jk - js + 1); }
  if (idir == -1) { int jj = je - js + 1; return 1 + 2 * jj - (je - jk + 1); }
  if (idir == 2) { int kk = ke - ks + 1; return kd - 2 * kk + (jk - ks + 1); }
  int kk = ke - ks + 1;
  return 1 + 2 * kk - (ke - jk + 1);
}

__device__ __forceinline__ double rho_of(const DevMesh& M, long long i) { return M.q[0][i] * M.q[M.nq - 1][i]; }

// findturm (src/komegasst.f:203-211) with X = F2*rho*strain*rei
__device__ __forceinline__ double findturm(double rho, double q5, double q6, double X) {
  const double a1 = 0.31;
  double t = a1 * rho * q5 / dmax(a1 * q6, X);
  return dmax(t, 1e-20);
}
}  // namespace

// komegabc for one instance, segments in the reference's order
__global__ __launch_bounds__(BC_THREADS) void k_sst_bc(const DevMesh* __restrict__ Ms, int eddy) {
  const DevMesh& M = Ms[blockIdx.x];
  if (*M.status) return;
  double* q5 = M.q[4];
  double* q6 = M.q[5];
  const double* u = M.sa_u;
  const double* v = M.sa_v;
  const double rei = 1. / M.c.rey;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int ib = 0; ib < M.nbc; ++ib) {
    const DevBC& b = M.bc[ib];
    int js = b.js, je = b.je, ks = b.ks, ke = b.ke, idir = b.idir;
    int iadd = idir >= 0 ? 1 : -1, iadir = idir >= 0 ? idir : -idir;
    int ty = b.type;
    double foso = 1.0;
    if (ty == 3 || ty == 4) {  // komegabc_extpt, one halo layer
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd, j2 = j1 + iadd;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j1, k), i2 = cell(M, j2, k);
          q5[i] = (1. + foso) * q5[i1] - foso * q5[i2];
          q6[i] = (1. + foso) * q6[i1] - foso * q6[i2];
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd, k2 = k1 + iadd;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j, k1), i2 = cell(M, j, k2);
          q5[i] = (1. + foso) * q5[i1] - foso * q5[i2];
          q6[i] = (1. + foso) * q6[i1] - foso * q6[i2];
        }
      }
    } else if (ty == 5 && iadir == 2) {  // komegabc_wall (src/komegasst.f:1034-1049)
      double tt = ((double)*M.istep0 - 1.) / 30.;
      if (tt > 1. || M.c.iread > 0) tt = 1.;
      double scal = (10. - 15. * tt + 6. * tt * tt) * ((tt * tt) * tt);
      int k = (idir == 2) ? ke : ks, k1 = k + iadd, kv = k1 - 1;
      for (int j = js + t; j <= je; j += nt) {
        int jv = j - 1;
        long long i = cell(M, j, k), i1 = cell(M, j, k1);
        double dely2 = sqk(M.x[i1] - 0.5 * (M.xv[vert(M, jv, kv)] + M.xv[vert(M, jv + 1, kv)])) +
                       sqk(M.y[i1] - 0.5 * (M.yv[vert(M, jv, kv)] + M.yv[vert(M, jv + 1, kv)]));
        double vm1 = M.sa_vmul[i1];
        double vmulwall = 0.5 * (vm1 + vm1);
        double omegawall = 60. * vmulwall * rei * rei / (KW_beta1 * dely2);
        double qwall5 = -q5[i1];
        double qwall6 = 2 * omegawall - 6. * vm1 * rei * rei / (KW_beta1 * dely2);
        q5[i] = scal * qwall5 + (1. - scal) * q5[i];
        q6[i] = scal * qwall6 + (1. - scal) * q6[i];
      }
    } else if (ty == 10 || ty == 11) {  // komegabc_sym
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = (idir == 1) ? je + 1 : js - 1;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j1, k);
          q5[i] = q5[i1]; q6[i] = q6[i1];
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j, k1);
          q5[i] = q5[i1]; q6[i] = q6[i1];
        }
      }
    } else if (ty == 22) {  // komegabc_periodic
      int nj = je - js + 1, nk = ke - ks + 1;
      for (int w = t; w < nj * nk; w += nt) {
        int j = js + w % nj, k = ks + w / nj;
        long long src = (iadir == 1) ? cell(M, kw_periodic_src(idir, js, je, ks, ke, M.jd, M.kd, j), k)
                                     : cell(M, j, kw_periodic_src(idir, js, je, ks, ke, M.jd, M.kd, k));
        long long i = cell(M, j, k);
        q5[i] = q5[src]; q6[i] = q6[src];
      }
    } else if (ty == 51 && iadir == 2) {  // komegabc_wake
      int k = (idir == 2) ? ke : ks, k1 = (idir == 2) ? ke + 1 : ks - 1;
      for (int j = js + t; j <= je; j += nt) {
        int jj = M.jd - j + 1;
        long long a = cell(M, jj, k1), c = cell(M, j, k1);
        double a5 = q5[a], a6 = q6[a], c5 = q5[c], c6 = q6[c];
        q5[cell(M, j, k)] = a5; q6[cell(M, j, k)] = a6;
        q5[cell(M, jj, k)] = c5; q6[cell(M, jj, k)] = c6;
      }
    } else if (ty == 46) {  // komegabc_inf
      int nj = je - js + 1, nk = ke - ks + 1;
      for (int w = t; w < nj * nk; w += nt) {
        long long i = cell(M, js + w % nj, ks + w / nj);
        double r = rho_of(M, i);
        q5[i] = M.c.tkeinf * r;
        q6[i] = M.c.tomegainf * r;
      }
    } else if (ty == 47) {  // komegabc_out
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd;
        for (int k = ks + t; k <= ke; k += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j1, k);
          double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
          double ug1 = M.has_gv ? M.ug[i1] : 0.0, vg1 = M.has_gv ? M.vg[i1] : 0.0;
          double r = rho_of(M, i);
          double v5 = M.c.tkeinf * r, v6 = M.c.tomegainf * r;
          double uu = (u[i] - ug + u[i1] - ug1) * M.xx[i];
          uu = uu + (v[i] - vg + v[i1] - vg1) * M.xy[i];
          uu = uu * iadd;
          if (uu < 0.) { v5 = q5[i1]; v6 = q6[i1]; }
          q5[i] = v5; q6[i] = v6;
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd;
        for (int j = js + t; j <= je; j += nt) {
          long long i = cell(M, j, k), i1 = cell(M, j, k1);
          double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
          double ug1 = M.has_gv ? M.ug[i1] : 0.0, vg1 = M.has_gv ? M.vg[i1] : 0.0;
          double r = rho_of(M, i);
          double v5 = M.c.tkeinf * r, v6 = M.c.tomegainf * r;
          double uu = (u[i] - ug + u[i1] - ug1) * M.yx[i];
          uu = uu + (v[i] - vg + v[i1] - vg1) * M.yy[i];
          uu = uu * iadd;
          if (uu < 0.) { v5 = q5[i1]; v6 = q6[i1]; }
          q5[i] = v5; q6[i] = v6;
        }
      }
    }
    __syncthreads();
  }
  if (!eddy) return;
  // findturm on the halo ring (F2 = strain = 0 there)
  const int jd = M.jd, kd = M.kd;
  int nring = 2 * jd + 2 * (kd - 2);
  for (int w = t; w < nring; w += nt) {
    int j, k;
    if (w < jd) { j = w + 1; k = 1; }
    else if (w < 2 * jd) { j = w - jd + 1; k = kd; }
    else if (w < 2 * jd + kd - 2) { j = 1; k = w - 2 * jd + 2; }
    else { j = jd; k = w - 2 * jd - (kd - 2) + 2; }
    long long i = cell(M, j, k);
    M.turmu[i] = findturm(rho_of(M, i), q5[i], q6[i], M.kw_x[i]);
  }
}

// kwvortic (strain), turbstress, crossdiff and blend for the interior cells,
// then findturm for every cell.  turbstress reads the eddy viscosity of the
// previous step at the cell itself, which this kernel then replaces.
__global__ __launch_bounds__(256) void k_sst_blend(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int jd = M.jd, js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const double rei = 1. / M.c.rey;
  const double* q5 = M.q[4];
  const double* q6 = M.q[5];
  const double* u = M.sa_u;
  const double* v = M.sa_v;
  long long n2 = (long long)jd * M.kd;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += 256ll * gridDim.x) {
    const int j = 1 + (int)(i % jd), k = 1 + (int)(i / jd);
    const double rho = rho_of(M, i);
    double f1 = 0.0, X = 0.0;
    if (j >= js && j <= je && k >= ks && k <= ke) {
      const long long ie = i + 1, iw = i - 1, in = i + jd, is = i - jd;
      const double xx = M.xx[i], xy = M.xy[i], yx = M.yx[i], yy = M.yy[i];
      double usi = 0.5 * (u[ie] - u[iw]), vsi = 0.5 * (v[ie] - v[iw]);
      double ueta = 0.5 * (u[in] - u[is]), veta = 0.5 * (v[in] - v[is]);
      double ux = xx * usi + yx * ueta, uy = xy * usi + yy * ueta;
      double vx = xx * vsi + yx * veta, vy = xy * vsi + yy * veta;
      double sxx = ux, sxy = 0.5 * (uy + vx), syy = vy;
      double strain = sqrt(2. * (sxx * sxx + 2. * sxy * sxy + syy * syy));
      double tm = M.turmu[i];
      double sfdiv = 2. / 3 * (ux + vy);
      double tauxx = tm * rei * (2.0 * ux - sfdiv) - 2. / 3 * q5[i];
      double tauyy = tm * rei * (2.0 * vy - sfdiv) - 2. / 3 * q5[i];
      double tauxy = tm * rei * (uy + vx);
      M.kw_prod[i] = tauxx * ux + tauxy * (uy + vx) + tauyy * vy;
      // crossdiff
      double re = rho_of(M, ie), rw = rho_of(M, iw), rn = rho_of(M, in), rs = rho_of(M, is);
      double ksi = 0.5 * (q5[ie] / re - q5[iw] / rw);
      double keta = 0.5 * (q5[in] / rn - q5[is] / rs);
      double omegasi = 0.5 * (q6[ie] / re - q6[iw] / rw);
      double omegaeta = 0.5 * (q6[in] / rn - q6[is] / rs);
      double kx = ksi * xx + keta * yx, ky = ksi * xy + keta * yy;
      double omegax = omegasi * xx + omegaeta * yx, omegay = omegasi * xy + omegaeta * yy;
      double cdkw = kx * omegax + ky * omegay;
      cdkw = cdkw * 2 * KW_sigmaw2 * rho * rho / q6[i];
      M.kw_cd[i] = cdkw;
      // blend
      double sn = M.sa_sn[i];
      double a1 = rei * sqrt(q5[i] * rho) / (0.09 * q6[i] * sn);
      double a2 = 500 * rei * rei * M.sa_vmul[i] / (q6[i] * sqk(sn));
      double cd = dmax(cdkw, 1e-10);
      double a3 = 4 * q5[i] * KW_sigmaw2 / (cd * sqk(sn));
      double arg1 = dmin(dmax(a1, a2), a3);
      double arg2 = dmax(2 * a1, a2);
      f1 = tanh(((arg1 * arg1) * arg1) * arg1);
      double f2 = tanh(arg2 * arg2);
      X = f2 * rho * strain * rei;
    }
    M.kw_f1[i] = f1;
    M.kw_x[i] = X;
    M.turmu[i] = findturm(rho, q5[i], q6[i], X);
  }
}

// j-face diffusion coefficient and difference of equation n between cells
// a and b = a+1 (src/komegasst.f:630-636, 676-682)
__device__ __forceinline__ void kw_half(const DevMesh& M, long long a, long long b, const double* qn, double s1,
                                        double s2, double rei, double& chp, double& dn) {
  double vnulh = 0.5 * (M.sa_vmul[a] + M.sa_vmul[b]);
  double vnuh = 0.5 * (M.turmu[a] + M.turmu[b]);
  double f1 = M.kw_f1[a];
  double sig = f1 * s1 + (1. - f1) * s2;
  chp = rei * (vnulh + sig * vnuh);
  dn = qn[b] / rho_of(M, b) - qn[a] / rho_of(M, a);
}

// komegarhslhs (convection, source, diffusion; src/komegasst.f:464-903) and
// the pseudo-time scaling (104-131) for every interior cell, each cell
// accumulating its own rows in the reference's order
__global__ __launch_bounds__(256) void k_sst_rhs(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const DevConst& C = M.c;
  const int jd = M.jd, js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const int nj = je - js + 1, nk = ke - ks + 1;
  const double rey = C.rey, rei = 1. / rey;
  const double* xx = M.xx;
  const double* xy = M.xy;
  const double* yx = M.yx;
  const double* yy = M.yy;
  long long ncell = (long long)nj * nk;
  for (long long w = blockIdx.x * 256ll + threadIdx.x; w < ncell; w += 256ll * gridDim.x) {
    const int j = js + (int)(w % nj), k = ks + (int)(w / nj);
    const long long i = cell(M, j, k);
    const long long iw = i - 1, ie = i + 1, is = i - jd, in = i + jd;
    const double rho = rho_of(M, i);
    double ug = M.has_gv ? M.ug[i] : 0.0, vg = M.has_gv ? M.vg[i] : 0.0;
    double uu = xx[i] * (M.sa_u[i] - ug) + xy[i] * (M.sa_v[i] - vg);
    double up = 0.5 * (uu + fabs(uu)), um = 0.5 * (uu - fabs(uu));
    double vv = yx[i] * (M.sa_u[i] - ug) + yy[i] * (M.sa_v[i] - vg);
    double vp = 0.5 * (vv + fabs(vv)), vm = 0.5 * (vv - fabs(vv));
    double sj[2], aj[2], bj[2], cj[2], ak[2], ck[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const double* qn = M.q[4 + n];
      double s = M.s[4 + n][i];
      s = s - up * (qn[i] - qn[iw]) - um * (qn[ie] - qn[i]);
      aj[n] = 0.0 - up;
      cj[n] = 0.0 + um;
      double b = 0.0 + up - um;
      s = s - vp * (qn[i] - qn[is]) - vm * (qn[in] - qn[i]);
      ak[n] = 0.0 - vp;
      ck[n] = 0.0 + vm;
      b = b + vp - vm;
      sj[n] = s;
      bj[n] = b;
    }
    // ---- komegasource (src/komegasst.f:867-901)
    {
      const double q5 = M.q[4][i], q6 = M.q[5][i];
      const double f1 = M.kw_f1[i];
      double prod = M.kw_prod[i];
      double kdes = KW_betastar * rey * q6 * q5 / rho;
      double ksust = KW_betastar * rey * C.tkeinf * C.tomegainf * rho;
      double prodlim = 10. * kdes;
      prod = dmin(prod, prodlim);
      double kprod = prod;
      double beta = f1 * KW_beta1 + (1 - f1) * KW_beta2;
      double omegades = beta * rey * q6 * q6 / rho;
      double omegasust = beta * rey * (C.tomegainf * C.tomegainf) * rho;
      double omegacrossdiff = (1 - f1) * M.kw_cd[i] * rei;
      sj[0] = sj[0] + kprod - kdes + ksust;
      double gammap = f1 * KW_gamma1 + (1 - f1) * KW_gamma2;
      sj[1] = sj[1] + gammap * rho / M.turmu[i] * prod - omegades + omegacrossdiff + omegasust;
      bj[0] = bj[0] + kdes / q5;
      bj[1] = bj[1] + (2 * omegades + fabs(omegacrossdiff)) / q6;
    }
    // ---- komegadiffus (src/komegasst.f:593-844)
    const double rw = rho_of(M, iw), re = rho_of(M, ie), rs = rho_of(M, is), rn = rho_of(M, in);
    {
      double dmxhm = 0.5 * (xx[iw] + xx[i]), dmyhm = 0.5 * (xy[iw] + xy[i]);
      double dmxhp = 0.5 * (xx[i] + xx[ie]), dmyhp = 0.5 * (xy[i] + xy[ie]);
      double dxp = dmxhp * xx[i] + dmyhp * xy[i];
      double dxm = dmxhm * xx[i] + dmyhm * xy[i];
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const double s1 = n == 0 ? KW_sigmak1 : KW_sigmaw1, s2 = n == 0 ? KW_sigmak2 : KW_sigmaw2;
        double chpm, dnm, chpp, dnp;
        kw_half(M, iw, i, M.q[4 + n], s1, s2, rei, chpm, dnm);
        kw_half(M, i, ie, M.q[4 + n], s1, s2, rei, chpp, dnp);
        double dcp = dxp * chpp, dcm = dxm * chpm;
        double ax = dmax(dcm, 0.0), cx = dmax(dcp, 0.0);
        sj[n] = sj[n] - ax * dnm + cx * dnp;
        aj[n] = aj[n] - ax / rw;
        cj[n] = cj[n] - cx / re;
        bj[n] = bj[n] + (ax + cx) / rho;
      }
    }
    {
      double dmxhm = 0.5 * (yx[is] + yx[i]), dmyhm = 0.5 * (yy[is] + yy[i]);
      double dmxhp = 0.5 * (yx[i] + yx[in]), dmyhp = 0.5 * (yy[i] + yy[in]);
      double dyp = dmxhp * yx[i] + dmyhp * yy[i];
      double dym = dmxhm * yx[i] + dmyhm * yy[i];
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const double s1 = n == 0 ? KW_sigmak1 : KW_sigmaw1, s2 = n == 0 ? KW_sigmak2 : KW_sigmaw2;
        double chpm, dnm, chpp, dnp;
        kw_half(M, is, i, M.q[4 + n], s1, s2, rei, chpm, dnm);
        kw_half(M, i, in, M.q[4 + n], s1, s2, rei, chpp, dnp);
        double dcp = dyp * chpp, dcm = dym * chpm;
        double ay = dmax(dcm, 0.0), cy = dmax(dcp, 0.0);
        sj[n] = sj[n] - ay * dnm + cy * dnp;
        ak[n] = ak[n] - ay / rs;
        ck[n] = ck[n] - cy / rn;
        bj[n] = bj[n] + (ay + cy) / rho;
      }
    }
    // ---- pseudo-time scaling (src/komegasst.f:100-118)
    double tscal = M.tscale[i];
    if (C.timeac == 1.) {
      double qj = M.q[M.nq - 1][i];
      const double dtpseudo_turb = 10.0, oat = time_oat(M);
      tscal = (double)imk(M.iblank[i], 0) * (1.0 + 0.002 * sqrt(qj)) / (1. + sqrt(qj));
      tscal = tscal * dtpseudo_turb;
      tscal = tscal / (1. + tscal / C.h / oat);
    }
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      M.kw_a[n][i] = aj[n] * tscal;
      M.kw_c[n][i] = cj[n] * tscal;
      M.kw_ak[n][i] = ak[n] * tscal;
      M.kw_ck[n][i] = ck[n] * tscal;
      M.kw_b[n][i] = 1. + bj[n] * tscal;
      M.kw_s[n][i] = sj[n] * tscal;
    }
  }
}

// lsolvej2 / lsolvek2 of one line held in LDS, eliminated from both ends
// (lanes 0 and 1; the reference's one-sided recurrence for lines of up to
// 32 cells); on return ls holds the solution
__device__ __forceinline__ void kw_line_solve(const double* la, const double* lb, const double* lc, double* ls,
                                              double* lcp, int n) {
  int nr, m;
  sa_split(n, nr, m);
  if (threadIdx.x >= (nr > 0 ? 2u : 1u)) return;
  thomas2_lds<1, 1>(la, lb, lc, ls, ls, lcp, n, (int)threadIdx.x);
}

// lsolvej2 for k row blockIdx.x / 2, equation blockIdx.x % 2, then
// sjmat = sjmat * bjmat (src/komegasst.f:133-140)
__global__ __launch_bounds__(64) void k_sst_solve_j(const DevMesh* __restrict__ Ms) {
  __shared__ double la[SA_LINE_MAX], lb[SA_LINE_MAX], lc[SA_LINE_MAX], ls[SA_LINE_MAX], lcp[SA_LINE_MAX];
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int eq = blockIdx.x & 1;
  const int k = M.ks + (blockIdx.x >> 1);
  if (k > M.ke) return;
  const int n = M.je - M.js + 1;
  const long long i0 = cell(M, M.js, k);
  double* __restrict__ s = M.kw_s[eq];
  const double* __restrict__ a = M.kw_a[eq];
  const double* __restrict__ b = M.kw_b[eq];
  const double* __restrict__ c = M.kw_c[eq];
  stage_line4(a, b, c, s, i0, 1, n, la, lb, lc, ls);
  __syncthreads();
  kw_line_solve(la, lb, lc, ls, lcp, n);
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += 64) s[i0 + t] = ls[t] * lb[t];
}

// lsolvek2 for j column blockIdx.x / 2, equation blockIdx.x % 2, then the
// update of that equation with its clip (src/komegasst.f:146-172)
__global__ __launch_bounds__(64) void k_sst_solve_k(const DevMesh* __restrict__ Ms) {
  __shared__ double la[SA_LINE_MAX], lb[SA_LINE_MAX], lc[SA_LINE_MAX], ls[SA_LINE_MAX], lcp[SA_LINE_MAX];
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int eq = blockIdx.x & 1;
  const int j = M.js + (blockIdx.x >> 1);
  if (j > M.je) return;
  const int n = M.ke - M.ks + 1;
  const long long i0 = cell(M, j, M.ks), st = M.jd;
  const double* __restrict__ s = M.kw_s[eq];
  const double* __restrict__ a = M.kw_ak[eq];
  const double* __restrict__ b = M.kw_b[eq];
  const double* __restrict__ c = M.kw_ck[eq];
  stage_line4(a, b, c, s, i0, st, n, la, lb, lc, ls);
  __syncthreads();
  kw_line_solve(la, lb, lc, ls, lcp, n);
  __syncthreads();
  double* qn = M.q[4 + eq];
  double* sn = M.s[4 + eq];
  for (int t = threadIdx.x; t < n; t += 64) {
    long long i = i0 + t * st;
    double ds = 1. * ls[t];
    sn[i] = ds;
    double q = qn[i] + ds * (double)imk(M.iblank[i], 0);
    qn[i] = dmax(q, 1e-20 * rho_of(M, i));
  }
}

// findturm for the interior after the update (the halo ring follows the
// final komegabc in k_sst_bc)
__global__ __launch_bounds__(256) void k_sst_turm(const DevMesh* __restrict__ Ms) {
  const DevMesh& M = Ms[blockIdx.y];
  if (*M.status) return;
  const int js = M.js, je = M.je, ks = M.ks, ke = M.ke;
  const int nj = je - js + 1, nk = ke - ks + 1;
  long long ncell = (long long)nj * nk;
  for (long long w = blockIdx.x * 256ll + threadIdx.x; w < ncell; w += 256ll * gridDim.x) {
    const long long i = cell(M, js + (int)(w % nj), ks + (int)(w / nj));
    M.turmu[i] = findturm(rho_of(M, i), M.q[4][i], M.q[5][i], M.kw_x[i]);
  }
}

static inline int nblk3(long long n, int per) {
  long long b = (n + per - 1) / per;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// komegasst for every instance (src/komegasst.f:2-187)
hipError_t launch_sst(const DevMesh* dM, const DevMesh* hM, int nmi, Marks* mk, hipStream_t st) {
  long long mx = 0;
  int mj = 1, mk_ = 1;
  for (int i = 0; i < nmi; ++i) {
    mx = hM[i].jd * (long long)hM[i].kd > mx ? hM[i].jd * (long long)hM[i].kd : mx;
    mj = hM[i].je - hM[i].js + 1 > mj ? hM[i].je - hM[i].js + 1 : mj;
    mk_ = hM[i].ke - hM[i].ks + 1 > mk_ ? hM[i].ke - hM[i].ks + 1 : mk_;
  }
  hipError_t e;
  if ((e = launch_sa_prim(dM, hM, nmi, st)) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_PRIM, st);
  hipLaunchKernelGGL(k_sst_bc, dim3(nmi), dim3(BC_THREADS), 0, st, dM, 0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_BC, st);
  hipLaunchKernelGGL(k_sst_blend, dim3(nblk3(mx, 256), nmi), dim3(256), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_sst_rhs, dim3(nblk3(mx, 256), nmi), dim3(256), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_RHS, st);
  hipLaunchKernelGGL(k_sst_solve_j, dim3(2 * mk_, nmi), dim3(64), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_SOLVE_J, st);
  hipLaunchKernelGGL(k_sst_solve_k, dim3(2 * mj, nmi), dim3(64), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_SOLVE_K, st);
  hipLaunchKernelGGL(k_sst_bc, dim3(nmi), dim3(BC_THREADS), 0, st, dM, 1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_sst_turm, dim3(nblk3(mx, 256), nmi), dim3(256), 0, st, dM);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  mark(mk, OT2D_ST_SA_EDDY, st);
  return hipSuccess;
}

}  // namespace ot2d
