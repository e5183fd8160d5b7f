// Overset interpolation between meshes (do_interpolations,
// src/connectivity.f:2-189), after every iteration of a case with several
// meshes.  Two passes, as the reference makes them: every donor first forms
// its 3x3 quadratic-Lagrange value from the current state (k_interp_donors),
// then every fringe cell takes the value of its donor (k_interp_fringes);
// a fringe cell can be a donor's neighbour, so the passes cannot be fused.
#include "ot2d_dev.h"

namespace ot2d {

// one donor: instance record, donor cell (1-based, solver indexing) and its
// fractions; one fringe: instance record, cell, 0-based donor row
struct InterpDonor { int mi, jj, kk, pad; double f1, f2; };
struct InterpFringe { int mi, jj, kk, src; };

__global__ __launch_bounds__(256) void k_interp_donors(const DevMesh* __restrict__ Ms,
                                                       const InterpDonor* __restrict__ don, int ndon,
                                                       double* __restrict__ qbc, int nv) {
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d >= ndon) return;
  const InterpDonor D = don[d];
  const DevMesh& M = Ms[D.mi];
  if (*M.status) return;
  const double onefourth = 1. / 4;
  double djm1 = 1. + D.f1, dkm1 = 1. + D.f2;
  double dj0 = D.f1, dk0 = D.f2;
  double djp1 = 1. - D.f1, dkp1 = 1. - D.f2;
  double w[9];
  w[0] = (dj0 * djp1 * dk0 * dkp1) * onefourth;
  w[1] = -2 * (djm1 * djp1 * dk0 * dkp1) * onefourth;
  w[2] = -(djm1 * dj0 * dk0 * dkp1) * onefourth;
  w[3] = -2 * (dj0 * djp1 * dkm1 * dkp1) * onefourth;
  w[4] = 4 * (djm1 * djp1 * dkm1 * dkp1) * onefourth;
  w[5] = 2 * (djm1 * dj0 * dkm1 * dkp1) * onefourth;
  w[6] = -(dj0 * djp1 * dkm1 * dk0) * onefourth;
  w[7] = 2 * (djm1 * djp1 * dkm1 * dk0) * onefourth;
  w[8] = (djm1 * dj0 * dkm1 * dk0) * onefourth;
  long long c[9];
  int s = 0;
  for (int dk = -1; dk <= 1; ++dk)
    for (int dj = -1; dj <= 1; ++dj) c[s++] = cell(M, D.jj + dj, D.kk + dk);
  const double* qj = M.q[M.nq - 1];
  double* out = qbc + (size_t)d * nv;
  // conserved variables (q times J) for n <= nmv, raw turbulence variables
  for (int n = 0; n < nv; ++n) {
    const double* q = M.q[n];
    double a;
    if (n < 4) {
      a = w[0] * q[c[0]] * qj[c[0]];
      for (int i = 1; i < 9; ++i) a = a + w[i] * q[c[i]] * qj[c[i]];
    } else {
      a = w[0] * q[c[0]];
      for (int i = 1; i < 9; ++i) a = a + w[i] * q[c[i]];
    }
    out[n] = a;
  }
}

__global__ __launch_bounds__(256) void k_interp_fringes(const DevMesh* __restrict__ Ms,
                                                        const InterpFringe* __restrict__ fr, int nfr,
                                                        const double* __restrict__ qbc, int nv) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nfr) return;
  const InterpFringe F = fr[f];
  const DevMesh& M = Ms[F.mi];
  if (*M.status) return;
  const long long i = cell(M, F.jj, F.kk);
  const double* v = qbc + (size_t)F.src * nv;
  const double qj = M.q[M.nq - 1][i];
  for (int n = 0; n < nv; ++n) M.q[n][i] = n < 4 ? v[n] / qj : v[n];
}

hipError_t launch_interp(const DevMesh* dM, const void* don, int ndon, const void* fr, int nfr, double* qbc, int nv,
                         hipStream_t st) {
  if (ndon > 0)
    hipLaunchKernelGGL(k_interp_donors, dim3((ndon + 255) / 256), dim3(256), 0, st, dM,
                       static_cast<const InterpDonor*>(don), ndon, qbc, nv);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (nfr > 0)
    hipLaunchKernelGGL(k_interp_fringes, dim3((nfr + 255) / 256), dim3(256), 0, st, dM,
                       static_cast<const InterpFringe*>(fr), nfr, qbc, nv);
  return hipGetLastError();
}

size_t interp_donor_bytes() { return sizeof(InterpDonor); }
size_t interp_fringe_bytes() { return sizeof(InterpFringe); }

}  // namespace ot2d
