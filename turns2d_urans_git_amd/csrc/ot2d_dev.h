// Device-side data structures shared by the gfx950 kernels and the C-ABI
// host code of libot2d.
//
// State layout in HBM is the reference's own (src/step.f:17-29): per mesh,
// Fortran column-major cell arrays (jd,kd) with j fastest, one halo ring,
// variables of q and s as separate planes q(jd,kd,nspec,nq).  A DevMesh
// record describes one (mesh, time instance) pair: base pointers already
// offset to the instance, so every kernel works on flat j-fastest planes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ot2d.h"

#define OT2D_MAXQ 8
// longest grid line (interior cells) of the turbulence-model line solves
#define SA_LINE_MAX 1024
#define OT2D_MAXBC_DEV 25

namespace ot2d {

// One boundary segment with the reference's index fix-ups already applied
// (src/bc.f:21-35): js..je, ks..ke are 1-based cell indices incl. halo.
struct DevBC {
  int type, idir, js, je, ks, ke;
  // segments of one phase touch disjoint cells and run side by side; the
  // phases run in order, so segments that share cells keep the reference's
  // segment order
  int phase;
};

// Case constants derived as read_inputs does (src/setup.f:1162-1248) plus
// the scheme switches the kernels branch on.
constexpr int OT2D_MAX_ITN = 16;   // Newton sub-iterations per time step, at most

struct DevConst {
  double gamma, gm1, ggm1, einf, pinf, rinf, uinf, vinf, rey, rmue, h;
  double th, qt, epse, Mp, dualtime, timeac, jac_factor, pr, tinf;
  double vnuinf, rotf;
  double tkeinf, tomegainf;  // SST free stream (src/setup.f:1232-1234)
  double angmax, xac, yac;   // time-spectral pitching (src/setup.f:348-393)
  double fmm, half, ca, sa;  // force normalisation (src/forces.f:27-31, 166-172)
  int invisc, iread, ilim, iconstant_cfl, idual, iprecon, timespectral, ntac;
  int iturb, nmv, motiontype;
  int itnmax;                // Newton sub-iterations per time step
  int iunst;                 // 1: the grid pitches every time step (motiontype 1)
  double dtps[OT2D_MAX_ITN]; // dtpseudo(itn), the dual time of sub-iteration itn
};

struct DevMesh {
  int jd, kd, jmax, kmax;
  int js, je, ks, ke;        // jbeg, jend, kbeg, kend (1-based)
  int nq, nv, nbc, nspec, nsp;
  int nbc_phase;             // phases of the boundary segments
  int jtail1, jtail2;
  int has_gv;                // grid velocities present (time-spectral)
  int res_block0, res_nblocks;  // this instance's slice of the norm partials
  double epsj, epsk;         // Koren limiter epsilons (src/eulerrhs.f:51-52)
  DevBC bc[OT2D_MAXBC_DEV];
  double* q[OT2D_MAXQ];      // q(:,:,nsp,n) planes, n = 1..nq  (index n-1)
  double* s[OT2D_MAXQ];      // s(:,:,nsp,n) planes
  double *x, *y, *xx, *xy, *yx, *yy, *ug, *vg, *turmu, *tscale, *bt;
  double* turmu_n;           // SA beside ilu2d: the new eddy viscosity until the update takes it
  double *xv, *yv, *ugv, *vgv;  // vertex arrays (jmax,kmax) of this instance
  const double *xv0, *yv0;      // instance-1 vertices (time-spectral rotation)
  double *xold, *yold, *xole, *yole;  // vertices one and two time steps back (moving grids), or null
  int* iblank;
  int bodyflag;              // mesh has a wall segment (type 5)
  // SA workspace (jd,kd) planes: laminar viscosity, u, v, wall distance and
  // the rows of the two tridiagonal systems (src/samodel2d.f:22-30)
  double *sa_vmul, *sa_u, *sa_v, *sa_sn, *sa_sj, *sa_aj, *sa_cj, *sa_ak, *sa_ck, *sa_bb;
  // SST workspace (src/komegasst.f:20-41): F1, F2*rho*strain/Re, production,
  // cross diffusion, and per equation (k, omega) the rows of the j and k
  // tridiagonal systems sharing one diagonal
  double *kw_f1, *kw_x, *kw_prod, *kw_cd;
  double *kw_s[2], *kw_a[2], *kw_b[2], *kw_c[2], *kw_ak[2], *kw_ck[2];
  double* lu;                // LU-SGS sweep records (workspace)
  int lu_rl;                 // rows per lane of the LU-SGS sweeps (1..4)
  int lu_fp32;               // sweep arithmetic: 1 FP32, 0 FP64
  // per-case bookkeeping (shared by all instances of a case)
  int* status;               // 0 ok, 1 mstop, 2 norm
  int* istep0;               // device copy of istep0
  int* istep;                // time steps of this run (the driver's istep; time-accurate steps)
  double* ta;                // time-accurate levels: qtn, qtnm1, qnewt (nv planes each), or null
  double* partials;          // norm partials of this instance's tiles (res_nblocks * 4 doubles)
  double* norms;             // rsum, resrho, resmax, jmax, kmax, nmax
  int* stepc;                // steps started in the current ot2d_iterate call
  double* hist;              // residual history rsum[step][slot] (may be null)
  int hist_cap, hist_stride, hist_slot;
  double* tspec;             // time instant of this instance
  const double* Ds;          // spectral derivative matrix (nspec x nspec)
  double* qall[OT2D_MAXQ];   // q planes of instance 0 (TS coupling base)
  long long inst_stride;     // elements between instances within a plane set
  DevConst c;
};

// Residual work item: one TJ x TK tile of interior cells of one instance.
struct ResTile {
  int mesh;   // index into the DevMesh array
  int j0, k0; // first interior cell (1-based)
  int blk;    // tile index within the instance (norm partial slot)
};

// Pointers into device memory as the global address space: loads through
// them are global (not flat) instructions, whose completions the compiler
// counts exactly, so loads issued ahead of their use stay in flight.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ gdouble* gptr(double* p) { return (gdouble*)p; }
__device__ __forceinline__ const gdouble* gptr(const double* p) { return (const gdouble*)p; }

// Stage one grid line of four coefficient arrays into LDS (element t at
// p[i0 + t*st]) with a 64-lane workgroup.  Each lane issues its loads for
// eight elements of all four arrays before its first LDS write, so a line
// of up to 512 elements arrives in one memory round trip rather than one
// per element and array.
__device__ __forceinline__ void stage_line4(const double* a, const double* b, const double* c, const double* s,
                                            long long i0, long long st, int n, double* la, double* lb, double* lc,
                                            double* ls) {
  for (int t0 = 0; t0 < n; t0 += 512) {
    double ra[8], rb[8], rc[8], rs[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + (int)threadIdx.x + 64 * u;
      const long long i = i0 + (long long)(t < n ? t : n - 1) * st;
      ra[u] = a[i]; rb[u] = b[i]; rc[u] = c[i]; rs[u] = s[i];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + (int)threadIdx.x + 64 * u;
      if (t < n) { la[t] = ra[u]; lb[t] = rb[u]; lc[t] = rc[u]; ls[t] = rs[u]; }
    }
  }
}

// Thomas recurrence of one tridiagonal line held in LDS, run by one thread
// with the reference's arithmetic (lsolvej / lsolvek and their two-equation
// forms):
//   forward   bb = 1/(b - a c'),  s' = (s - a s') bb,  c' = c bb
//   backward  s' = s' - c' s'_next
// s' and c' go to os and oc (os may be ls itself).  The recurrence runs in
// chunks of THOMAS_CH cells whose LDS reads are issued a chunk ahead into
// registers, so the dependent chain does not wait on LDS.  Chunk reads past
// the line's end are clamped to its last cell and never used.  Cell t of the
// line sits at index t*ST of each array (ST > 1: several lines interleaved,
// one per lane).
constexpr int THOMAS_CH = 8;
template <int ST = 1>
__device__ __forceinline__ void thomas_lds(const double* la, const double* lb, const double* lc, const double* ls,
                                           double* os, double* oc, int n) {
  constexpr int CH = THOMAS_CH;
  double bb = 1. / lb[0];
  double sp = ls[0] * bb, cp = lc[0] * bb;
  os[0] = sp; oc[0] = cp;
  double pa[CH], pb[CH], pc[CH], ps[CH], qa[CH], qb[CH], qc[CH], qs[CH];
  auto ld = [&](double (&xa)[CH], double (&xb)[CH], double (&xc)[CH], double (&xs)[CH], int t) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int i = (t + u < n ? t + u : n - 1) * ST;
      xa[u] = la[i]; xb[u] = lb[i]; xc[u] = lc[i]; xs[u] = ls[i];
    }
  };
  auto run = [&](const double (&xa)[CH], const double (&xb)[CH], const double (&xc)[CH], const double (&xs)[CH],
                 int t, int cnt) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (u < cnt) {
        const double ai = xa[u];
        bb = 1. / (xb[u] - ai * cp);
        sp = (xs[u] - ai * sp) * bb;
        cp = xc[u] * bb;
        os[(t + u) * ST] = sp; oc[(t + u) * ST] = cp;
      }
    }
  };
  int t = 1;
  ld(pa, pb, pc, ps, t);
  for (; t + 2 * CH <= n; t += 2 * CH) {
    ld(qa, qb, qc, qs, t + CH);
    run(pa, pb, pc, ps, t, CH);
    ld(pa, pb, pc, ps, t + 2 * CH);
    run(qa, qb, qc, qs, t + CH, CH);
  }
  ld(qa, qb, qc, qs, t + CH);
  run(pa, pb, pc, ps, t, n - t);
  run(qa, qb, qc, qs, t + CH, n - t - CH);
  // back substitution over cells n-2 .. 0, chunks of cells t-CH+1 .. t
  double sn = sp;
  double ys[CH], yc[CH], zs[CH], zc[CH];
  auto ldb = [&](double (&xs)[CH], double (&xc)[CH], int t) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int i = (t - u >= 0 ? t - u : 0) * ST;
      xs[u] = os[i]; xc[u] = oc[i];
    }
  };
  auto runb = [&](const double (&xs)[CH], const double (&xc)[CH], int t, int cnt) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (u < cnt) {
        sn = xs[u] - xc[u] * sn;
        os[(t - u) * ST] = sn;
      }
    }
  };
  int b = n - 2;
  ldb(ys, yc, b);
  for (; b - 2 * CH + 1 >= 0; b -= 2 * CH) {
    ldb(zs, zc, b - CH);
    runb(ys, yc, b, CH);
    ldb(ys, yc, b - 2 * CH);
    runb(zs, zc, b - CH, CH);
  }
  ldb(zs, zc, b - CH);
  runb(ys, yc, b, b + 1);
  runb(zs, zc, b - CH, b - CH + 1);
}

// Two-ended elimination of a tridiagonal line (the turbulence models' line
// solves): with 32-cell tiles, ntile = ceil(n/32) and nr = floor(ntile/2)
// (0 for ntile = 1), the right part is the last 32*nr cells and the left
// part the first m = n - 32*nr:
//   left  (t = 0 .. m-1, up):   bb = 1/(b - a c'),  s' = (s - a s') bb,  c' = c bb
//   right (t = n-1 .. m, down): bb = 1/(b - c a'),  s' = (s - c s') bb,  a' = a bb
// (the first cell of each side: bb = 1/b); at the meeting cells each side
// takes x = (s' - c' s'_other) / (1 - c' c'_other) and substitutes back
// towards its end, x = s' - c' x_next.  nr = 0 is the one-sided recurrence.
constexpr int SAJ_TILE = 32;
__device__ __forceinline__ void sa_split(int n, int& nr, int& m) {
  const int ntile = (n + SAJ_TILE - 1) / SAJ_TILE;
  nr = ntile >= 2 ? ntile / 2 : 0;
  m = n - SAJ_TILE * nr;
}

// The two-ended elimination of one line in LDS, run by two lanes of one
// wave, h = 0 (left) and h = 1 (right), lane numbers XM apart (nr > 0), or
// by the h = 0 lane alone (nr = 0).  Cell t of the line sits at index t*ST.
// s' and the solution go to os (which may be ls), c' / a' to oc.
template <int ST, int XM>
__device__ __forceinline__ void thomas2_lds(const double* la, const double* lb, const double* lc, const double* ls,
                                            double* os, double* oc, int n, int h) {
  constexpr int CH = THOMAS_CH;
  int nr, m;
  sa_split(n, nr, m);
  const int cnt = h ? n - m : m;
  const int i0 = (h ? n - 1 : 0) * ST, st = (h ? -1 : 1) * ST;   // index of step u: i0 + st*u
  const double* lo = h ? lc : la;
  const double* hi = h ? la : lc;
  double bb = 1. / lb[i0];
  double sp = ls[i0] * bb, cp = hi[i0] * bb;
  os[i0] = sp; oc[i0] = cp;
  double pa[CH], pb[CH], pc[CH], ps[CH], qa[CH], qb[CH], qc[CH], qs[CH];
  auto ld = [&](double (&xa)[CH], double (&xb)[CH], double (&xc)[CH], double (&xs)[CH], int u0) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int i = i0 + st * (u0 + u < cnt ? u0 + u : cnt - 1);
      xa[u] = lo[i]; xb[u] = lb[i]; xc[u] = hi[i]; xs[u] = ls[i];
    }
  };
  auto run = [&](const double (&xa)[CH], const double (&xb)[CH], const double (&xc)[CH], const double (&xs)[CH],
                 int u0, int k) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (u < k) {
        const double ai = xa[u];
        bb = 1. / (xb[u] - ai * cp);
        sp = (xs[u] - ai * sp) * bb;
        cp = xc[u] * bb;
        const int i = i0 + st * (u0 + u);
        os[i] = sp; oc[i] = cp;
      }
    }
  };
  int u = 1;
  ld(pa, pb, pc, ps, u);
  for (; u + 2 * CH <= cnt; u += 2 * CH) {
    ld(qa, qb, qc, qs, u + CH);
    run(pa, pb, pc, ps, u, CH);
    ld(pa, pb, pc, ps, u + 2 * CH);
    run(qa, qb, qc, qs, u + CH, CH);
  }
  ld(qa, qb, qc, qs, u + CH);
  run(pa, pb, pc, ps, u, cnt - u);
  run(qa, qb, qc, qs, u + CH, cnt - u - CH);
  // the meeting cells
  double sn = sp;
  if (nr > 0) {
    const double so = __shfl_xor(sp, XM), co = __shfl_xor(cp, XM);
    sn = (sp - cp * so) / (1. - cp * co);
  }
  os[i0 + st * (cnt - 1)] = sn;
  // back substitution towards this side's end, chunks of CH steps
  double ys[CH], yc[CH], zs[CH], zc[CH];
  auto ldb = [&](double (&xs)[CH], double (&xc)[CH], int u1) {
#pragma unroll
    for (int v = 0; v < CH; ++v) {
      const int i = i0 + st * (u1 - v >= 0 ? u1 - v : 0);
      xs[v] = os[i]; xc[v] = oc[i];
    }
  };
  auto runb = [&](const double (&xs)[CH], const double (&xc)[CH], int u1, int k) {
#pragma unroll
    for (int v = 0; v < CH; ++v) {
      if (v < k) {
        sn = xs[v] - xc[v] * sn;
        os[i0 + st * (u1 - v)] = sn;
      }
    }
  };
  int b = cnt - 2;
  ldb(ys, yc, b);
  for (; b - 2 * CH + 1 >= 0; b -= 2 * CH) {
    ldb(zs, zc, b - CH);
    runb(ys, yc, b, CH);
    ldb(ys, yc, b - 2 * CH);
    runb(zs, zc, b - CH, CH);
  }
  ldb(zs, zc, b - CH);
  runb(ys, yc, b, b + 1);
  runb(zs, zc, b - CH, b - CH + 1);
}

constexpr int RES_TJ = 32;
constexpr int RES_TK = 8;
constexpr int RES_THREADS = RES_TJ * RES_TK;

// boundary-condition kernels run one workgroup per instance, its segments
// one after another
constexpr int BC_THREADS = 256;

// oat of the time-accuracy order (src/time.f:22-25): the reference's istep
// counts the time steps of the run
__device__ inline double time_oat(const DevMesh& M) {
  double oat = 1.0;
  if (M.c.ntac < 2) return oat;
  const int istep = *M.istep;
  if (M.c.ntac >= 2 && istep > 1) oat = 2. / 3.;
  if (M.c.ntac == 3 && istep > 2) oat = 6. / 11.;
  return oat;
}

// round trip of ilu2d's in-place primitive conversion (src/lhs.f:524-531,
// 805-811): the reference leaves q(j,k,1:4) with these bits
__device__ __forceinline__ void q_roundtrip(double& q1, double& q2, double& q3, double& q4, double qj) {
  double dj = 1.0 / q1;
  double u = q2 * dj, v = q3 * dj, e = q4 * dj;
  double r = q1 * qj;
  q1 = r / qj;
  q2 = u * q1; q3 = v * q1; q4 = e * q1;
}

// laminar viscosity (lamvis, src/samodel2d.f:200-238) and the velocities
// vmu_sa keeps (src/samodel2d.f:61-72) of one cell from its state
__device__ inline void sa_prim_cell(const DevMesh& M, long long i, double q1, double q2, double q3, double q4) {
  const double c2b = 198.6 / M.c.tinf, c2bp = c2b + 1.;
  double ra = 1. / q1;
  double uvel = q2 * ra, vvel = q3 * ra;
  double ei = q4 * ra - 0.5 * (uvel * uvel + vvel * vvel);
  double tt = M.c.ggm1 * ei;
  M.sa_vmul[i] = c2bp * tt * sqrt(tt) / (c2b + tt);
  M.sa_u[i] = q2 / q1;
  M.sa_v[i] = q3 / q1;
}

__host__ __device__ inline long long cell(const DevMesh& M, int j, int k) {
  return (long long)(j - 1) + (long long)(k - 1) * M.jd;
}

// residual norms of step (src/step.f:53-115), first stage: the partial sums
// (sum of squares, rho part, max |s| with its first (k,n,j) position) of the
// cells of one tile, reduced in a fixed order and stored per tile
__device__ inline void norm_tile(const DevMesh& M, const ResTile& T, int j, int k, bool valid, const double* s,
                                 double* redbuf) {
  double(*red)[RES_THREADS] = reinterpret_cast<double(*)[RES_THREADS]>(redbuf);
  const int tid = threadIdx.x;
  const int jd = M.jd;
  double rs = 0.0, rr1 = 0.0, rmax = 0.0, key = 0.0;
  if (valid) {
    long long i = cell(M, j, k);
    double ibp = (double)(M.iblank[i] > 0 ? M.iblank[i] : 0);
    double m1 = s[0] * ibp, m2 = s[1] * ibp, m3 = s[2] * ibp, m4 = s[3] * ibp;
    rs = m1 * m1 + m2 * m2 + m3 * m3 + m4 * m4;
    rr1 = s[0] * s[0];
    // first maximum in (k, n, j) order (src/step.f:67-78)
    for (int n = 0; n < 4; ++n) {
      double a = fabs(s[n]) * ibp;
      if (a > rmax) { rmax = a; key = (double)(((long long)(k - 1) * 4 + n) * jd + (j - 1)); }
    }
  }
  red[0][tid] = rs; red[1][tid] = rr1; red[2][tid] = rmax; red[3][tid] = valid && rmax > 0.0 ? key : 1e300;
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
    double* P = M.partials + 4ll * T.blk;
    P[0] = red[0][0]; P[1] = red[1][0]; P[2] = red[2][0]; P[3] = red[3][0];
  }
}
__host__ __device__ inline long long vert(const DevMesh& M, int j, int k) {
  return (long long)(j - 1) + (long long)(k - 1) * M.jmax;
}

}  // namespace ot2d

// launch wrappers (ot2d_kernels.hip)
namespace ot2d {
// Optional per-kernel HIP events for profiling: mark(m, s, st) is recorded on
// the launch stream right after the kernel of stage s (an OT2D_ST_* value).
struct Marks {
  static constexpr int CAP = 64;
  hipEvent_t ev[CAP];
  int stage[CAP];
  int n = 0;
};
inline void mark(Marks* m, int s, hipStream_t st) {
  if (m && m->n < Marks::CAP) {
    (void)hipEventRecord(m->ev[m->n], st);
    m->stage[m->n++] = s;
  }
}

// metrics_only: metfv alone on moved vertices (no cell centres, no
// free-stream state), as the time loop of a moving grid does
hipError_t launch_init_geometry(const DevMesh* dM, int nmi, hipStream_t st, int metrics_only = 0);
// pitch (src/motion.f:546-632): the grid turns by this time step's angle
hipError_t launch_pitch(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
// with ncase > 0 also opens the step of every running case (istep0 and the
// step counter)
hipError_t launch_time_step(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st, int* istep0 = nullptr,
                            const int* status = nullptr, int* stepc = nullptr, int ncase = 0, int itn = 0);
// time-accurate steps with Newton sub-iterations (src/time.f:109-209)
hipError_t launch_open_step(const int* status, int* istep0, int* stepc, int* istep, int ncase, hipStream_t st);
hipError_t launch_stqol(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
hipError_t launch_newton(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
hipError_t launch_copy_doubles(double* dst, const double* src, int n, hipStream_t st);
hipError_t launch_bc(const DevMesh* dM, const DevMesh* hM, int nmi, int zero_s, hipStream_t st);
hipError_t launch_residual(const DevMesh* dM, const ResTile* tiles, int ntiles, int do_norms, int weno,
                           hipStream_t st);
hipError_t launch_norms(const DevMesh* dM, const ResTile* tiles, int ntiles, hipStream_t st);
// rbc: the residual boundary conditions follow in the same launch
hipError_t launch_norms_final(const DevMesh* dM, int nmi, hipStream_t st, int rbc = 0);
hipError_t launch_conv_check(const DevMesh* dM, const int* inst, const double* tol, int* status, int ncase,
                             hipStream_t st);
hipError_t launch_zero_s(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
hipError_t launch_rbc(const DevMesh* dM, int nmi, hipStream_t st);
// launch choices that change no result (the autotuner's knobs)
constexpr int LU32_SWEEP_VARIANTS = 3;   // FP32 sweep instantiations (ring steps, producers, prefetch)
struct LaunchCfg {
  int sweep = 0;     // 0: 12 ring steps, 3 producers; 1: 18, 3; 2: 12, 4 (with deeper prefetch)
  int sa_lane = -1;  // SA line solves: -1 by the number of lines, 0 one wave per line, 1 one lane per line
};
// fuse_update: 1 the q update and monitor in k_lu_finish (Euler), | 2
// k_lu_finish also leaves the SA model's laminar viscosity and velocities
// (what k_sa_prim computes) from the state it writes
hipError_t launch_lu(const DevMesh* dM, const DevMesh* hM, int nmi, int fuse_update, int fp32, Marks* mk,
                     hipStream_t st, int sweep = 0);
hipError_t launch_monitor(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
hipError_t launch_update_monitor(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st, int take_turmu = 0);
// k_lu_finish alone (launch_lu with fuse_update | 4 leaves it out)
hipError_t launch_lu_finish(const DevMesh* dM, const DevMesh* hM, int nmi, int flags, int fp32, hipStream_t st);
hipError_t launch_forces(const DevMesh* dM, int mi, double* out, hipStream_t st);
hipError_t launch_tsrhs(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
size_t lu_workspace_bytes(int jd, int kd, int rl);
hipError_t launch_visrhs(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st, int norms = 0);
// prim_done: the laminar viscosity and velocities are already in place
// (k_lu_finish with flag 2)
// ovl: the SA model runs beside ilu2d (another stream, before k_lu_finish):
// it takes q1-q4 through ilu2d's primitive round trip itself and leaves the
// eddy viscosity in turmu_n, which k_update_monitor copies
hipError_t launch_sa(const DevMesh* dM, const DevMesh* hM, int nmi, Marks* mk, hipStream_t st, int sa_lane = -1,
                     int prim_done = 0, int ovl = 0);
hipError_t launch_sa_dist(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
hipError_t launch_sa_prim(const DevMesh* dM, const DevMesh* hM, int nmi, hipStream_t st);
hipError_t launch_sst(const DevMesh* dM, const DevMesh* hM, int nmi, Marks* mk, hipStream_t st);
// overset interpolation (ot2d_overset.hip): don/fr are device arrays of
// {int mi, jj, kk, pad; double f1, f2} and {int mi, jj, kk, src}
hipError_t launch_interp(const DevMesh* dM, const void* don, int ndon, const void* fr, int nfr, double* qbc, int nv,
                         hipStream_t st);
size_t interp_donor_bytes();
size_t interp_fringe_bytes();
size_t lu_lds_need(int ni, int fp32);
}  // namespace ot2d
