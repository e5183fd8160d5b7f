// Host-side implicit hole cutting (ot2d_ihc.cpp): inputs and outputs of one
// connectivity pass over a set of overlapping meshes.
#pragma once
#include <cstdint>
#include <vector>

namespace ot2d {
namespace ihc {

// One mesh as connect2d sees it: vertex coordinates (jmax x kmax, j
// fastest), cell centres and cell areas of the cells 0..jmax x 0..kmax
// (halo ring included; the solver's cell (j+1,k+1)), and the rows of its
// ihcbc file.
struct MeshIn {
  int jmax, kmax;
  const double *xv, *yv;     // (jmax, kmax)
  const double *xc, *yc;     // (jmax+1, kmax+1)
  const double* vol;         // (jmax+1, kmax+1)
  int nbc;
  int ibtyp[25], jbcs[25], jbce[25], kbcs[25], kbce[25];
  double XV(int j, int k) const { return xv[(size_t)(j - 1) + (size_t)(k - 1) * jmax]; }
  double YV(int j, int k) const { return yv[(size_t)(j - 1) + (size_t)(k - 1) * jmax]; }
  double XC(int j, int k) const { return xc[(size_t)j + (size_t)k * (jmax + 1)]; }
  double YC(int j, int k) const { return yc[(size_t)j + (size_t)k * (jmax + 1)]; }
  double VOL(int j, int k) const { return vol[(size_t)j + (size_t)k * (jmax + 1)]; }
};

// Lists of one mesh in the reference's layout (imesh/idonor pairs already
// shifted to solver cell indices, 1-based; ibc indexes the donor lists of
// all meshes concatenated, 1-based).
struct MeshOut {
  std::vector<int> iblank;   // solver cells (jmax+1, kmax+1) before the halo copy
  std::vector<int> imesh;    // 2 * nfringe: fringe cell (j,k)
  std::vector<int> ibc;      // nfringe
  std::vector<int> idonor;   // 2 * ndonor: donor cell (j,k) of this mesh
  std::vector<double> frac;  // 2 * ndonor
  int iisptr = 1, iieptr = 0;
};

struct Result {
  std::vector<MeshOut> mesh;
};

// connect2d over ngrids meshes; returns 0, or -1 for meshes too small
int connect(const MeshIn* meshes, int ngrids, int nhalo, Result* out);

}  // namespace ihc
}  // namespace ot2d
