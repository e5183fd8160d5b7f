/*
 * ot2d_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C restatement of the reference OVERTURNS-2D iteration on host
 * memory.  Every routine names the reference file:line it follows.  Operation
 * order follows the Fortran expressions left to right (no contraction:
 * built with -ffp-contract=off) so that the restatement agrees with the
 * reference built by oracle/Makefile bit for bit where the reference's own
 * arithmetic is deterministic; tests check that against oracle/_ref dumps and
 * the committed fixtures under tests/golden/.
 *
 * Layout: Fortran column-major, 1-based indices in the macros below.
 */
#include "ot2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

typedef struct or_mesh {
  int jmax, kmax, jd, kd, jbeg, jend, kbeg, kend;
  int jtail1, jtail2, jle, jtail_v;
  int bodyflag;
  double rotf, xac, yac;
  ot2d_mesh_desc d;
  size_t n2;  /* jd*kd */
  size_t nv2; /* jmax*kmax */
  double *q, *s, *x, *y, *xv, *yv, *xx, *xy, *yx, *yy, *ug, *vg, *ugv, *vgv;
  double *turmu, *tscale, *bt;
  int32_t* iblank;
  /* overset lists per time instance (reference layout, 1-based indices) */
  int32_t *nfringe, *ndonor;
  int32_t **imesh, **ibc, **idonor;
  double** frac;
} or_mesh;

struct or_case {
  ot2d_params p;
  /* derived (read_inputs, src/setup.f:1162-1248) */
  double gm1, ggm1, einf, pinf, rinf, ainf, htinf, uinf, vinf, rey, h, pi;
  double rmue, tkeinf, tomegainf;
  int lamin, iturb, nq, nv, nmv, nspec, timespectral, iunst;
  int nmesh, istep0, istep, mdim;
  double* Ds;
  double* tspec;
  or_mesh* m;
  int overset;   /* connectivity installed: interpolate after every step */
};

/* ---------------------------------------------------------------- access */
#define IX(M, j, k) ((size_t)((j)-1) + (size_t)((k)-1) * (size_t)(M)->jd)
#define QA(M, nsp, n) ((M)->q + ((size_t)((n)-1) * c->nspec + (size_t)(nsp)) * (M)->n2)
#define SA(M, nsp, n) ((M)->s + ((size_t)((n)-1) * c->nspec + (size_t)(nsp)) * (M)->n2)
#define GA(M, arr, nsp) ((M)->arr + (size_t)(nsp) * (M)->n2)
#define VA(M, arr, nsp) ((M)->arr + (size_t)(nsp) * (M)->nv2)
#define VX(M, j, k) ((size_t)((j)-1) + (size_t)((k)-1) * (size_t)(M)->jmax)

static double sq(double a) { return a * a; }
static double cube(double a) { return a * a * a; }
/* Fortran SIGN(a,b) */
static double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------- defaults */
int32_t or_default_params(ot2d_params* p) {
  /* data statements, src/modules.f:170-193 */
  memset(p, 0, sizeof(*p));
  p->gamma = 1.4; p->pr = 0.72; p->rmue = 1.0; p->tinf = 400.0;
  p->iread = 0; p->nhalo = 1; p->half = 0;
  p->nsteps = 500; p->nrest = 1000; p->npnorm = 25;
  p->timespectral = 0; p->nspec = 1;
  p->fsmach = 0.16; p->fmtip = 0.0; p->alfa = 0.0; p->rartio = 1.0; p->rey = 3900000.0;
  p->invisc = 1; p->lamin = 1; p->iturb = 1; p->itrans = 0;
  p->vnuinf = 0.1; p->tuinf = 0.2;
  p->iunst = 0; p->ntac = 1; p->itnmax = 1; p->dt = 0.1; p->timeac = 1.0; p->motiontype = 1;
  p->epse = 0.01; p->irhsy = -3; p->ilhs = 1; p->ilim = 0; p->totime = 0.0; p->rf = 0.0;
  p->angmax = 0.0;
  p->idual = 0; p->dualtime = 1.0; p->iconstant_cfl = 0;
  p->iprecon = 0; p->Mp = 1.0;
  return OT2D_OK;
}

int32_t or_sizes(const ot2d_params* p, const ot2d_mesh_desc* m, int32_t* jd, int32_t* kd,
                 int32_t* nq, int32_t* nv) {
  int nadd = 2 * p->nhalo - 1; /* src/setup.f:1263 */
  int iturb = p->invisc ? -1 : (p->lamin ? -1 : p->iturb);
  int q = 5, v = 4;
  if (iturb == 1) { q += 1; v += 1; }
  if (iturb == 2) { q += 2; v += 2; }
  if ((iturb == 1 || iturb == 2) && p->itrans == 1) { q += 2; v += 2; }
  if (jd) *jd = m->jmax + nadd;
  if (kd) *kd = m->kmax + nadd;
  if (nq) *nq = q;
  if (nv) *nv = v;
  return OT2D_OK;
}

/* ---------------------------------------------------------------- create */
int32_t or_case_create(const ot2d_params* p, int32_t nmesh, const ot2d_mesh_desc* md,
                       or_case** out) {
  if (!p || !md || nmesh < 1 || !out) return OT2D_ERR_ARG;
  if (p->nhalo != 1) return OT2D_ERR_UNSUPPORTED;
  or_case* c = (or_case*)calloc(1, sizeof(or_case));
  c->p = *p;
  /* read_inputs, src/setup.f:963-1265 */
  c->timespectral = p->timespectral;
  c->iunst = p->timespectral ? 0 : p->iunst;
  c->nspec = p->timespectral ? p->nspec : 1;
  c->rmue = p->rmue;
  c->lamin = p->lamin;
  c->iturb = p->iturb;
  if (p->invisc) { c->rmue = 0.0; c->lamin = 1; }
  if (c->lamin) c->iturb = -1;
  c->pi = 4.0 * atan(1.0);
  c->gm1 = p->gamma - 1.0;
  c->ggm1 = p->gamma * c->gm1;
  double cs = cos(c->pi * p->alfa / 180.0);
  double ss = sin(c->pi * p->alfa / 180.0);
  c->einf = 1.0 / c->ggm1 + 0.5 * sq(p->fsmach);
  c->pinf = 1.0 / p->gamma;
  c->rinf = 1.0;
  c->ainf = p->gamma * c->pinf / c->rinf;
  c->htinf = p->gamma * c->einf / c->rinf - 0.5 * c->gm1 * sq(p->fsmach);
  c->nq = 5; c->nv = 4; c->nmv = 4;
  if (c->iturb == 1) { c->nq += 1; c->nv += 1; }
  else if (c->iturb == 2) { c->nq += 2; c->nv += 2; }
  if ((c->iturb == 1 || c->iturb == 2) && p->itrans == 1) { c->nq += 2; c->nv += 2; }
  c->uinf = p->fsmach * cs;
  c->vinf = p->fsmach * ss;
  c->rey = (p->fsmach != 0.0) ? p->rey / p->fsmach : p->rey / p->fmtip;
  c->h = p->dt;
  if (p->itrans == 0) {
    c->tkeinf = 1e-6 * sq(p->fsmach);
    c->tomegainf = 5 * p->fsmach / c->rey;
  } else {
    double tu = p->tuinf;
    c->tkeinf = sq(tu / 100 * p->fsmach) * 1.5;
    c->tomegainf = c->tkeinf / p->vnuinf;
  }
  c->nmesh = nmesh;
  c->mdim = 0;
  c->m = (or_mesh*)calloc((size_t)nmesh, sizeof(or_mesh));
  for (int im = 0; im < nmesh; ++im) {
    or_mesh* M = &c->m[im];
    M->d = md[im];
    M->jmax = md[im].jmax;
    M->kmax = md[im].kmax;
    M->jd = M->jmax + 1;
    M->kd = M->kmax + 1;
    c->mdim = MAX(c->mdim, MAX(M->jd, M->kd));
    /* set_pointers_globals, src/setup.f:55-108 */
    M->jbeg = 2; M->kbeg = 2; M->jend = M->jd - 1; M->kend = M->kd - 1;
    M->bodyflag = 0;
    M->jtail_v = 1; /* initia, src/init.f:59-66 */
    for (int ib = 0; ib < md[im].nbc; ++ib)
      if (md[im].bc[ib].ibtyp == 5) { M->bodyflag = 1; M->jtail_v = md[im].bc[ib].jbcs; }
    M->jtail1 = M->jtail_v + 1;
    if (p->half == 0) { M->jle = (M->jd + 1) / 2; M->jtail2 = M->jd - M->jtail1 + 1; }
    else { M->jle = M->jd - 1; M->jtail2 = M->jd - 1; }
    M->rotf = (p->rf != 0.0) ? md[im].irot_dir * p->rf : 0.0;
    M->xac = md[im].xcen; M->yac = md[im].ycen;
    M->n2 = (size_t)M->jd * M->kd;
    M->nv2 = (size_t)M->jmax * M->kmax;
    size_t ns = (size_t)c->nspec;
    M->q = (double*)calloc(M->n2 * ns * c->nq, sizeof(double));
    M->s = (double*)calloc(M->n2 * ns * c->nv, sizeof(double));
#define AL(a) M->a = (double*)calloc(M->n2 * ns, sizeof(double))
    AL(x); AL(y); AL(xx); AL(xy); AL(yx); AL(yy); AL(ug); AL(vg); AL(turmu); AL(tscale); AL(bt);
#undef AL
    M->xv = (double*)calloc(M->nv2 * ns, sizeof(double));
    M->yv = (double*)calloc(M->nv2 * ns, sizeof(double));
    M->ugv = (double*)calloc(M->nv2 * ns, sizeof(double));
    M->vgv = (double*)calloc(M->nv2 * ns, sizeof(double));
    M->iblank = (int32_t*)calloc(M->n2 * ns, sizeof(int32_t));
    for (size_t i = 0; i < M->n2 * ns; ++i) { M->iblank[i] = 1; M->bt[i] = 1.0; }
  }
  /* computeTScoefs, src/setup.f:838-873 */
  int N = c->nspec;
  c->Ds = (double*)calloc((size_t)N * N, sizeof(double));
  c->tspec = (double*)calloc((size_t)N, sizeof(double));
  for (int i = 1; i <= N; ++i)
    for (int j = 1; j <= N; ++j) {
      double v = 0.0;
      if (i != j) {
        double sgn = ((i - j) % 2 == 0) ? 1.0 : -1.0;
        if (N % 2 == 0) v = 0.5 * sgn / tan(c->pi * (i - j) / N);
        else v = 0.5 * sgn / sin(c->pi * (i - j) / N);
      }
      c->Ds[(i - 1) + (size_t)(j - 1) * N] = v;
    }
  *out = c;
  return OT2D_OK;
}

int32_t or_case_destroy(or_case* c) {
  if (!c) return OT2D_OK;
  for (int im = 0; im < c->nmesh; ++im) {
    or_mesh* M = &c->m[im];
    free(M->q); free(M->s); free(M->x); free(M->y); free(M->xx); free(M->xy); free(M->yx);
    free(M->yy); free(M->ug); free(M->vg); free(M->turmu); free(M->tscale); free(M->bt);
    free(M->xv); free(M->yv); free(M->ugv); free(M->vgv); free(M->iblank);
    if (M->nfringe) {
      for (int nsp = 0; nsp < c->nspec; ++nsp) {
        free(M->imesh[nsp]); free(M->ibc[nsp]); free(M->idonor[nsp]); free(M->frac[nsp]);
      }
      free(M->nfringe); free(M->ndonor); free(M->imesh); free(M->ibc); free(M->idonor); free(M->frac);
    }
  }
  free(c->m); free(c->Ds); free(c->tspec); free(c);
  return OT2D_OK;
}

/* BC index fix-ups, src/bc.f:21-35 */
static void bc_range(const or_mesh* M, const ot2d_bcseg* b, int* js, int* je, int* ks, int* ke) {
  int a = b->jbcs, e = b->jbce, f = b->kbcs, g = b->kbce;
  if (a < 0) a = M->jmax + a + 1;
  if (f < 0) f = M->kmax + f + 1;
  if (e < 0) e = M->jmax + e + 1;
  if (g < 0) g = M->kmax + g + 1;
  if (a > 1) a = a + 1;
  if (f > 1) f = f + 1;
  if (e < M->jmax) e = e + 1 - 1;
  if (g < M->kmax) g = g + 1 - 1;
  if (e == M->jmax) e = M->jd;
  if (g == M->kmax) g = M->kd;
  *js = a; *je = e; *ks = f; *ke = g;
}

/* ------------------------------------------------------ geometry (a14) */
/* find_cellcenter, src/setup.f:111-345 */
static void find_cellcenter(or_case* c, or_mesh* M, int nsp) {
  double* x = GA(M, x, nsp);
  double* y = GA(M, y, nsp);
  const double* xv = VA(M, xv, nsp);
  const double* yv = VA(M, yv, nsp);
  int jd = M->jd, kd = M->kd;
  (void)kd;
#define XC(j, k) x[IX(M, j, k)]
#define YC(j, k) y[IX(M, j, k)]
#define XV(j, k) xv[VX(M, j, k)]
#define YV(j, k) yv[VX(M, j, k)]
  memset(x, 0, M->n2 * sizeof(double));
  memset(y, 0, M->n2 * sizeof(double));
  for (int j = M->jbeg; j <= M->jend; ++j)
    for (int k = M->kbeg; k <= M->kend; ++k) {
      int jv = j, kv = k;
      XC(j, k) = 0.25 * (XV(jv, kv) + XV(jv - 1, kv) + XV(jv, kv - 1) + XV(jv - 1, kv - 1));
      YC(j, k) = 0.25 * (YV(jv, kv) + YV(jv - 1, kv) + YV(jv, kv - 1) + YV(jv - 1, kv - 1));
    }
  int j = M->jbeg - 1;
  for (int k = M->kbeg; k <= M->kend; ++k) {
    int jv = j, kv = k;
    XC(j, k) = XC(j + 1, k) - 0.5 * (XV(jv + 1, kv) - XV(jv, kv) + XV(jv + 1, kv - 1) - XV(jv, kv - 1));
    YC(j, k) = YC(j + 1, k) - 0.5 * (YV(jv + 1, kv) - YV(jv, kv) + YV(jv + 1, kv - 1) - YV(jv, kv - 1));
  }
  j = M->jend + 1;
  for (int k = M->kbeg; k <= M->kend; ++k) {
    int jv = j, kv = k;
    XC(j, k) = XC(j - 1, k) - 0.5 * (XV(jv - 2, kv) - XV(jv - 1, kv) + XV(jv - 2, kv - 1) - XV(jv - 1, kv - 1));
    YC(j, k) = YC(j - 1, k) - 0.5 * (YV(jv - 2, kv) - YV(jv - 1, kv) + YV(jv - 2, kv - 1) - YV(jv - 1, kv - 1));
  }
  int k = M->kbeg - 1;
  for (j = M->jbeg; j <= M->jend; ++j) {
    int jv = j, kv = k;
    XC(j, k) = XC(j, k + 1) - 0.5 * (XV(jv, kv + 1) - XV(jv, kv) + XV(jv - 1, kv + 1) - XV(jv - 1, kv));
    YC(j, k) = YC(j, k + 1) - 0.5 * (YV(jv, kv + 1) - YV(jv, kv) + YV(jv - 1, kv + 1) - YV(jv - 1, kv));
  }
  k = M->kend + 1;
  for (j = M->jbeg; j <= M->jend; ++j) {
    int jv = j, kv = k;
    XC(j, k) = XC(j, k - 1) - 0.5 * (XV(jv, kv - 2) - XV(jv, kv - 1) + XV(jv - 1, kv - 2) - XV(jv - 1, kv - 1));
    YC(j, k) = YC(j, k - 1) - 0.5 * (YV(jv, kv - 2) - YV(jv, kv - 1) + YV(jv - 1, kv - 2) - YV(jv - 1, kv - 1));
  }
  j = M->jbeg - 1; k = M->kbeg - 1;
  XC(j, k) = 2 * XC(j + 1, k) - XC(j + 2, k);
  YC(j, k) = 2 * YC(j + 1, k) - YC(j + 2, k);
  k = M->kend + 1;
  XC(j, k) = 2 * XC(j + 1, k) - XC(j + 2, k);
  YC(j, k) = 2 * YC(j + 1, k) - YC(j + 2, k);
  j = M->jend + 1; k = M->kbeg - 1;
  XC(j, k) = 2 * XC(j, k + 1) - XC(j, k + 2);
  YC(j, k) = 2 * YC(j, k + 1) - YC(j, k + 2);
  k = M->kend + 1;
  XC(j, k) = 2 * XC(j, k - 1) - XC(j, k - 2);
  YC(j, k) = 2 * YC(j, k - 1) - YC(j, k - 2);
  for (int ib = 0; ib < M->d.nbc; ++ib) {
    const ot2d_bcseg* b = &M->d.bc[ib];
    int js, je, ks, ke;
    bc_range(M, b, &js, &je, &ks, &ke);
    int idir = b->ibdir, iadir = abs(idir);
    if (b->ibtyp == 22) {
      if (idir == 1) {
        int jj = je - js + 1;
        for (j = js; j <= je; ++j) {
          int jc = jd - 2 * jj + (j - js + 1);
          for (k = ks; k <= ke; ++k) { XC(j, k) = XC(jc, k); YC(j, k) = YC(jc, k); }
        }
      } else if (idir == -1) {
        int jj = je - js + 1;
        for (j = js; j <= je; ++j) {
          int jc = 1 + 2 * jj - (je - j + 1);
          for (k = ks; k <= ke; ++k) { XC(j, k) = XC(jc, k); YC(j, k) = YC(jc, k); }
        }
      } else if (idir == 2) {
        int kk = ke - ks + 1;
        for (k = ks; k <= ke; ++k) {
          int kc = M->kd - 2 * kk + (k - ks + 1);
          for (j = js; j <= je; ++j) { XC(j, k) = XC(j, kc); YC(j, k) = YC(j, kc); }
        }
      } else if (idir == -2) {
        int kk = ke - ks + 1;
        for (k = ks; k <= ke; ++k) {
          int kc = 1 + 2 * kk - (ke - k + 1);
          for (j = js; j <= je; ++j) { XC(j, k) = XC(j, kc); YC(j, k) = YC(j, kc); }
        }
      }
    } else if (b->ibtyp == 51 && iadir == 2) {
      for (int kc = 1; kc <= ke - ks + 1; ++kc) {
        int kk, k1;
        if (idir == 2) { kk = ke - kc + 1; k1 = ke + kc; }
        else { kk = ks + kc - 1; k1 = ks - kc; }
        for (j = js; j <= je; ++j) {
          int jj = jd - j + 1;
          XC(j, kk) = XC(jj, k1); YC(j, kk) = YC(jj, k1);
          XC(jj, kk) = XC(j, k1); YC(jj, kk) = YC(j, k1);
        }
      }
    }
  }
#undef XC
#undef YC
}

/* metfv, src/setup.f:440-724 */
static void metfv(or_case* c, or_mesh* M, int nsp) {
  double* qj = QA(M, nsp, c->nq);
  double* xx = GA(M, xx, nsp);
  double* xy = GA(M, xy, nsp);
  double* yx = GA(M, yx, nsp);
  double* yy = GA(M, yy, nsp);
  const double* xv = VA(M, xv, nsp);
  const double* yv = VA(M, yv, nsp);
  int jd = M->jd;
  for (int k = 2; k <= M->kmax; ++k)
    for (int j = 2; j <= M->jmax; ++j) {
      double dx1 = XV(j, k) - XV(j - 1, k - 1);
      double dy1 = YV(j, k) - YV(j - 1, k - 1);
      double dx2 = XV(j - 1, k) - XV(j, k - 1);
      double dy2 = YV(j - 1, k) - YV(j, k - 1);
      double a3 = 0.5 * (dx1 * dy2 - dx2 * dy1);
      qj[IX(M, j, k)] = 1. / a3;
    }
  for (int j = 2; j <= M->jmax; ++j)
    for (int k = 2; k <= M->kmax; ++k) {
      xx[IX(M, j, k)] = 0.5 * (YV(j, k) - YV(j, k - 1) + YV(j - 1, k) - YV(j - 1, k - 1));
      xy[IX(M, j, k)] = -0.5 * (XV(j, k) - XV(j, k - 1) + XV(j - 1, k) - XV(j - 1, k - 1));
    }
  for (int k = 2; k <= M->kmax; ++k)
    for (int j = 2; j <= M->jmax; ++j) {
      yx[IX(M, j, k)] = -0.5 * (YV(j, k) - YV(j - 1, k) + YV(j, k - 1) - YV(j - 1, k - 1));
      yy[IX(M, j, k)] = 0.5 * (XV(j, k) - XV(j - 1, k) + XV(j, k - 1) - XV(j - 1, k - 1));
    }
  for (int k = M->kbeg; k <= M->kend; ++k)
    for (int j = M->jbeg; j <= M->jend; ++j) {
      size_t i = IX(M, j, k);
      double J = qj[i];
      xx[i] = J * xx[i]; xy[i] = J * xy[i]; yx[i] = J * yx[i]; yy[i] = J * yy[i];
    }
#define CP(dst, src) do { qj[dst] = qj[src]; xx[dst] = xx[src]; xy[dst] = xy[src]; \
    yx[dst] = yx[src]; yy[dst] = yy[src]; } while (0)
  {
    int j = M->jbeg - 1, j1 = M->jbeg;
    for (int k = M->kbeg; k <= M->kend; ++k) CP(IX(M, j, k), IX(M, j1, k));
    j = M->jend + 1; j1 = M->jend;
    for (int k = M->kbeg; k <= M->kend; ++k) CP(IX(M, j, k), IX(M, j1, k));
    int k = M->kbeg - 1, k1 = M->kbeg;
    for (j = M->jbeg - 1; j <= M->jend + 1; ++j) CP(IX(M, j, k), IX(M, j, k1));
    k = M->kend + 1; k1 = M->kend;
    for (j = M->jbeg - 1; j <= M->jend + 1; ++j) CP(IX(M, j, k), IX(M, j, k1));
  }
  for (int ib = 0; ib < M->d.nbc; ++ib) {
    const ot2d_bcseg* b = &M->d.bc[ib];
    int js, je, ks, ke;
    bc_range(M, b, &js, &je, &ks, &ke);
    int idir = b->ibdir, iadir = abs(idir);
    if (b->ibtyp == 22) {
      if (idir == 1) {
        int jj = je - js + 1;
        for (int j = js; j <= je; ++j) {
          int jc = jd - 2 * jj + (j - js + 1);
          for (int k = ks; k <= ke; ++k) CP(IX(M, j, k), IX(M, jc, k));
        }
      } else if (idir == -1) {
        int jj = je - js + 1;
        for (int j = js; j <= je; ++j) {
          int jc = 1 + 2 * jj - (je - j + 1);
          for (int k = ks; k <= ke; ++k) CP(IX(M, j, k), IX(M, jc, k));
        }
      } else if (idir == 2) {
        int kk = ke - ks + 1;
        for (int k = ks; k <= ke; ++k) {
          int kc = M->kd - 2 * kk + (k - ks + 1);
          for (int j = js; j <= je; ++j) CP(IX(M, j, k), IX(M, j, kc));
        }
      } else if (idir == -2) {
        int kk = ke - ks + 1;
        for (int k = ks; k <= ke; ++k) {
          int kc = 1 + 2 * kk - (ke - k + 1);
          for (int j = js; j <= je; ++j) CP(IX(M, j, k), IX(M, j, kc));
        }
      }
    } else if (b->ibtyp == 51 && iadir == 2) {
      for (int kc = 1; kc <= ke - ks + 1; ++kc) {
        int k, k1;
        if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; }
        else { k = ks + kc - 1; k1 = ks - kc; }
        for (int j = js; j <= je; ++j) {
          int jj = jd - j + 1;
          size_t a = IX(M, j, k), bb = IX(M, jj, k1), cc = IX(M, jj, k), dd = IX(M, j, k1);
          qj[a] = qj[bb]; xx[a] = -xx[bb]; xy[a] = -xy[bb]; yx[a] = -yx[bb]; yy[a] = -yy[bb];
          qj[cc] = qj[dd]; xx[cc] = -xx[dd]; xy[cc] = -xy[dd]; yx[cc] = -yx[dd]; yy[cc] = -yy[dd];
        }
      }
    }
  }
#undef CP
}
#undef XV
#undef YV

/* find_coord_TS, src/setup.f:348-393 (motiontype 1 and 2) */
static void find_coord_TS(or_case* c, or_mesh* M, int nsp) {
  const double* xv1 = VA(M, xv, 0);
  const double* yv1 = VA(M, yv, 0);
  double* xv = VA(M, xv, nsp);
  double* yv = VA(M, yv, nsp);
  double tspec = c->tspec[nsp];
  if (c->p.motiontype == 1) {
    double theta = c->p.angmax * c->pi / 180. * (1. - cos(M->rotf * tspec));
    double ss = sin(theta), cs = cos(theta);
    for (size_t i = 0; i < M->nv2; ++i) {
      xv[i] = (xv1[i] - M->xac) * cs + (yv1[i] - M->yac) * ss + M->xac;
      yv[i] = (yv1[i] - M->yac) * cs - (xv1[i] - M->xac) * ss + M->yac;
    }
  } else if (c->p.motiontype == 2) {
    double cs = cos(M->rotf * tspec), ss = sin(M->rotf * tspec);
    for (size_t i = 0; i < M->nv2; ++i) {
      xv[i] = (xv1[i] - M->xac) * cs - (yv1[i] - M->yac) * ss + M->xac;
      yv[i] = (yv1[i] - M->yac) * cs + (xv1[i] - M->xac) * ss + M->yac;
    }
  }
}

/* find_timemetrics_TS, src/setup.f:727-786 (motiontype 1 and 2) */
static void find_timemetrics_TS(or_case* c, or_mesh* M, int nsp) {
  double tspec = c->tspec[nsp];
  double* ug = GA(M, ug, nsp);
  double* vg = GA(M, vg, nsp);
  const double* x = GA(M, x, nsp);
  const double* y = GA(M, y, nsp);
  double* ugv = VA(M, ugv, nsp);
  double* vgv = VA(M, vgv, nsp);
  const double* xv = VA(M, xv, nsp);
  const double* yv = VA(M, yv, nsp);
  double rotf = M->rotf, pi = c->pi, am = c->p.angmax;
  if (c->p.motiontype == 1) {
    for (size_t i = 0; i < M->n2; ++i) {
      ug[i] = am * pi / 180. * rotf * (y[i] - M->yac) * sin(rotf * tspec);
      vg[i] = -am * pi / 180. * rotf * (x[i] - M->xac) * sin(rotf * tspec);
    }
    for (size_t i = 0; i < M->nv2; ++i) {
      ugv[i] = am * pi / 180. * rotf * (yv[i] - M->yac) * sin(rotf * tspec);
      vgv[i] = -am * pi / 180. * rotf * (xv[i] - M->xac) * sin(rotf * tspec);
    }
  } else if (c->p.motiontype == 2) {
    for (size_t i = 0; i < M->n2; ++i) { ug[i] = -rotf * (y[i] - M->yac); vg[i] = rotf * (x[i] - M->xac); }
    for (size_t i = 0; i < M->nv2; ++i) { ugv[i] = -rotf * (yv[i] - M->yac); vgv[i] = rotf * (xv[i] - M->xac); }
  }
}

/* qzero + qdivj, src/init.f:209-277, 180-206 */
static void qzero_qdivj(or_case* c, or_mesh* M, int nsp) {
  double ruinf = c->rinf * c->uinf, rvinf = c->rinf * c->vinf;
  double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
  double* qj = QA(M, nsp, c->nq);
  for (size_t i = 0; i < M->n2; ++i) { q1[i] = c->rinf; q2[i] = ruinf; q3[i] = rvinf; q4[i] = c->einf; }
  if (c->iturb == 1) {
    double* q5 = QA(M, nsp, 5);
    for (size_t i = 0; i < M->n2; ++i) q5[i] = c->p.vnuinf;
  } else if (c->iturb == 2) {
    double *q5 = QA(M, nsp, 5), *q6 = QA(M, nsp, 6);
    for (size_t i = 0; i < M->n2; ++i) { q5[i] = c->rinf * c->tkeinf; q6[i] = c->rinf * c->tomegainf; }
  }
  for (size_t i = 0; i < M->n2; ++i) {
    q1[i] = q1[i] / qj[i]; q2[i] = q2[i] / qj[i]; q3[i] = q3[i] / qj[i]; q4[i] = q4[i] / qj[i];
  }
}

/* initia, src/init.f:2-177 (iread=0 part) */
int32_t or_set_grid(or_case* c, int32_t im, const double* xv, const double* yv) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  or_mesh* M = &c->m[im];
  size_t ns = (size_t)c->nspec;
  for (size_t i = 0; i < M->n2 * ns; ++i) { M->turmu[i] = 0.; M->ug[i] = 0.; M->vg[i] = 0.; M->iblank[i] = 1; }
  memset(M->s, 0, M->n2 * ns * c->nv * sizeof(double));
  c->istep0 = 0;
  for (int nsp = 0; nsp < c->nspec; ++nsp)
    c->tspec[nsp] = (c->p.rf != 0.0) ? 2 * c->pi * nsp / c->nspec / fabs(c->p.rf) : 0.0;
  memcpy(M->xv, xv, M->nv2 * sizeof(double));
  memcpy(M->yv, yv, M->nv2 * sizeof(double));
  for (int nsp = 1; nsp < c->nspec; ++nsp) find_coord_TS(c, M, nsp);
  for (int nsp = 0; nsp < c->nspec; ++nsp) {
    find_cellcenter(c, M, nsp);
    metfv(c, M, nsp);
    if (c->timespectral) find_timemetrics_TS(c, M, nsp);
    qzero_qdivj(c, M, nsp);
  }
  return OT2D_OK;
}

/* ------------------------------------------------------------ state I/O */
int32_t or_set_state(or_case* c, int32_t im, const double* q) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  or_mesh* M = &c->m[im];
  memcpy(M->q, q, M->n2 * c->nspec * c->nq * sizeof(double));
  return OT2D_OK;
}
int32_t or_get_state(or_case* c, int32_t im, double* q) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  or_mesh* M = &c->m[im];
  memcpy(q, M->q, M->n2 * c->nspec * c->nq * sizeof(double));
  return OT2D_OK;
}
int32_t or_set_iblank(or_case* c, int32_t im, const int32_t* ib) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  memcpy(c->m[im].iblank, ib, c->m[im].n2 * c->nspec * sizeof(int32_t));
  return OT2D_OK;
}
int32_t or_get_iblank(or_case* c, int32_t im, int32_t* ib) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  memcpy(ib, c->m[im].iblank, c->m[im].n2 * c->nspec * sizeof(int32_t));
  return OT2D_OK;
}
static double* field_ptr(or_case* c, or_mesh* M, int field, size_t* n) {
  size_t ns = (size_t)c->nspec;
  *n = M->n2 * ns;
  switch (field) {
    case OT2D_F_X: return M->x;
    case OT2D_F_Y: return M->y;
    case OT2D_F_XX: return M->xx;
    case OT2D_F_XY: return M->xy;
    case OT2D_F_YX: return M->yx;
    case OT2D_F_YY: return M->yy;
    case OT2D_F_UG: return M->ug;
    case OT2D_F_VG: return M->vg;
    case OT2D_F_TSCALE: return M->tscale;
    case OT2D_F_BT: return M->bt;
    case OT2D_F_TURMU: return M->turmu;
    case OT2D_F_S: *n = M->n2 * ns * c->nv; return M->s;
    case OT2D_F_XV: *n = M->nv2 * ns; return M->xv;
    case OT2D_F_YV: *n = M->nv2 * ns; return M->yv;
    case OT2D_F_UGV: *n = M->nv2 * ns; return M->ugv;
    case OT2D_F_VGV: *n = M->nv2 * ns; return M->vgv;
  }
  return NULL;
}
int32_t or_get_field(or_case* c, int32_t im, int32_t field, double* out) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  size_t n;
  double* p = field_ptr(c, &c->m[im], field, &n);
  if (!p) return OT2D_ERR_ARG;
  memcpy(out, p, n * sizeof(double));
  return OT2D_OK;
}
int32_t or_set_field(or_case* c, int32_t im, int32_t field, const double* in) {
  if (!c || im < 0 || im >= c->nmesh) return OT2D_ERR_ARG;
  size_t n;
  double* p = field_ptr(c, &c->m[im], field, &n);
  if (!p) return OT2D_ERR_ARG;
  memcpy(p, in, n * sizeof(double));
  return OT2D_OK;
}
int32_t or_set_istep0(or_case* c, int32_t istep0) { c->istep0 = istep0; return OT2D_OK; }
int32_t or_get_istep0(or_case* c, int32_t* istep0) { *istep0 = c->istep0; return OT2D_OK; }

/* ------------------------------------------------------ time step (a8) */
/* time_step, src/time.f:2-107 */
static void time_step_mesh(or_case* c, or_mesh* M) {
  const ot2d_params* p = &c->p;
  double oat = 1.0;
  if (p->ntac >= 2 && c->istep > 1) oat = 2. / 3.;
  if (p->ntac == 3 && c->istep > 2) oat = 6. / 11.;
  double jac_factor = p->invisc ? 0.01 : 0.002;
  for (int nsp = 0; nsp < c->nspec; ++nsp) {
    double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
    double* qj = QA(M, nsp, c->nq);
    double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
    double *ug = GA(M, ug, nsp), *vg = GA(M, vg, nsp), *ts = GA(M, tscale, nsp), *bt = GA(M, bt, nsp);
    const int32_t* ib = M->iblank + (size_t)nsp * M->n2;
    for (size_t i = 0; i < M->n2; ++i) {
      double ibp = (double)MAX(ib[i], 0);
      if (p->iconstant_cfl == 0) {
        if (p->idual == 0) {
          ts[i] = oat * c->h * ibp * (1.0 + jac_factor * (1. - p->timeac) * sqrt(qj[i])) /
                  (1.0 + (1. - p->timeac) * sqrt(qj[i]));
          bt[i] = 1.;
        } else {
          double t = ibp * (1.0 + jac_factor * sqrt(qj[i])) / (1.0 + sqrt(qj[i]));
          if (p->iprecon) t = 100. * ibp * (1.0 + 0.005 * sqrt(qj[i])) / (1.0 + sqrt(qj[i]));
          t = t * p->dualtime;
          bt[i] = 1. / (1. + t / c->h / oat);
          ts[i] = t / (1. + t / c->h / oat);
        }
      } else {
        double u = q2[i] / q1[i], v = q3[i] / q1[i], e = q4[i] / q1[i];
        double uu = (u - ug[i]) * xx[i] + (v - vg[i]) * xy[i];
        double vv = (u - ug[i]) * yx[i] + (v - vg[i]) * yy[i];
        double dxdx = sq(xx[i]) + sq(xy[i]);
        double dydy = sq(yx[i]) + sq(yy[i]);
        double asq = c->ggm1 * (e - 0.5 * (sq(u) + sq(v)));
        double lam_j = fabs(uu) + sqrt(asq * dxdx);
        double lam_k = fabs(vv) + sqrt(asq * dydy);
        if (p->idual == 0) {
          ts[i] = c->h / (lam_j + lam_k);
          ts[i] = ts[i] * ibp;
          bt[i] = 1.;
        } else {
          double hnew = p->dualtime / (lam_j + lam_k);
          ts[i] = hnew * ibp;
          ts[i] = ts[i] / (1. + ts[i] / c->h / oat);
          bt[i] = 1. / (1. + ts[i] / c->h / oat);
        }
      }
    }
  }
}

int32_t or_time_step(or_case* c) {
  for (int im = 0; im < c->nmesh; ++im) time_step_mesh(c, &c->m[im]);
  return OT2D_OK;
}

/* ------------------------------------------------------------- BCs (a6) */
typedef struct {
  double *q1, *q2, *q3, *q4, *qj, *xx, *xy, *yx, *yy, *ug, *vg, *bt;
} qview;

static qview view(or_case* c, or_mesh* M, int nsp) {
  qview v;
  v.q1 = QA(M, nsp, 1); v.q2 = QA(M, nsp, 2); v.q3 = QA(M, nsp, 3); v.q4 = QA(M, nsp, 4);
  v.qj = QA(M, nsp, c->nq);
  v.xx = GA(M, xx, nsp); v.xy = GA(M, xy, nsp); v.yx = GA(M, yx, nsp); v.yy = GA(M, yy, nsp);
  v.ug = GA(M, ug, nsp); v.vg = GA(M, vg, nsp); v.bt = GA(M, bt, nsp);
  return v;
}

/* bcwall + bctany + uv, src/bc.f:1669-1801, 1328-1415, 2224-2264 */
static void bcwall(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir, int invsc) {
  int iadd = idir >= 0 ? 1 : -1, iadir = abs(idir);
  if (iadir != 2) return; /* "not implemented in bcwall" */
  int k = (idir == 2) ? ke : ks;
  int k1 = k + iadd, k2 = k1 + iadd;
  double foso = invsc ? 1.0 : 0.0;
  double gm1 = c->gm1;
  for (int j = js; j <= je; ++j) {
    size_t i = IX(M, j, k), i1 = IX(M, j, k1), i2 = IX(M, j, k2);
    double u, vv;
    /* bctany */
    if (invsc) {
      double f = 1.0;
      double u1 = (v->q2[i1] * v->xx[i] + v->q3[i1] * v->xy[i]) / v->q1[i1];
      double u2 = (v->q2[i2] * v->xx[i] + v->q3[i2] * v->xy[i]) / v->q1[i2];
      u = (1. + f) * u1 - f * u2;
      double v1 = (v->q2[i1] * v->yx[i] + v->q3[i1] * v->yy[i]) / v->q1[i1];
      vv = v->ug[i] * v->yx[i] + v->vg[i] * v->yy[i] + v->ug[i1] * v->yx[i1] + v->vg[i1] * v->yy[i1] - v1;
    } else {
      double u1 = (v->q2[i1] * v->xx[i] + v->q3[i1] * v->xy[i]) / v->q1[i1];
      double v1 = (v->q2[i1] * v->yx[i] + v->q3[i1] * v->yy[i]) / v->q1[i1];
      u = v->ug[i] * v->xx[i] + v->vg[i] * v->xy[i] + v->ug[i1] * v->xx[i1] + v->vg[i1] * v->xy[i1] - u1;
      vv = v->ug[i] * v->yx[i] + v->vg[i] * v->yy[i] + v->ug[i1] * v->yx[i1] + v->vg[i1] * v->yy[i1] - v1;
    }
    /* uv */
    double t11 = v->yy[i], t12 = -v->xy[i], t21 = -v->yx[i], t22 = v->xx[i];
    double vol = 1. / (t22 * t11 - t12 * t21);
    double uwall = (t11 * u + t12 * vv) * vol;
    double vwall = (t21 * u + t22 * vv) * vol;
    /* density / pressure extrapolation */
    double rho1 = v->q1[i1] * v->qj[i1];
    double rho2 = v->q1[i2] * v->qj[i2];
    double p1 = gm1 * (v->q4[i1] - 0.5 * (sq(v->q2[i1]) + sq(v->q3[i1])) / v->q1[i1]) * v->qj[i1];
    double p2 = gm1 * (v->q4[i2] - 0.5 * (sq(v->q2[i2]) + sq(v->q3[i2])) / v->q1[i2]) * v->qj[i2];
    v->q1[i] = ((1. + foso) * rho1 - foso * rho2) / v->qj[i];
    v->q2[i] = uwall * v->q1[i];
    v->q3[i] = vwall * v->q1[i];
    double press = (1. + foso) * p1 - foso * p2;
    v->q4[i] = press / (gm1 * v->qj[i]) + 0.5 * (sq(v->q2[i]) + sq(v->q3[i])) / v->q1[i];
  }
  /* other halo cells: ke-ks = 0 with one halo ring */
  for (int kc = 1; kc <= ke - ks; ++kc) {
    int kk = (idir == 2) ? ke - kc : ks + kc;
    int kk1 = kk + iadd, kk2 = kk1 + iadd;
    double f = 1.0;
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, kk), i1 = IX(M, j, kk1), i2 = IX(M, j, kk2);
      double rho1 = v->q1[i1] * v->qj[i1], rho2 = v->q1[i2] * v->qj[i2];
      double u1 = v->q2[i1] / v->q1[i1], u2 = v->q2[i2] / v->q1[i2];
      double v1 = v->q3[i1] / v->q1[i1], v2 = v->q3[i2] / v->q1[i2];
      double p1 = gm1 * (v->q4[i1] - 0.5 * (sq(v->q2[i1]) + sq(v->q3[i1])) / v->q1[i1]) * v->qj[i1];
      double p2 = gm1 * (v->q4[i2] - 0.5 * (sq(v->q2[i2]) + sq(v->q3[i2])) / v->q1[i2]) * v->qj[i2];
      v->q1[i] = ((1. + f) * rho1 - f * rho2) / v->qj[i];
      v->q2[i] = ((1. + f) * u1 - f * u2) * v->q1[i];
      v->q3[i] = ((1. + f) * v1 - f * v2) * v->q1[i];
      double press = (1. + f) * p1 - f * p2;
      v->q4[i] = press / (gm1 * v->qj[i]) + 0.5 * (sq(v->q2[i]) + sq(v->q3[i])) / v->q1[i];
    }
  }
  /* 30-step ramp, src/bc.f:1779-1793 */
  double t = (double)c->istep0 / 30.;
  if (t > 1. || c->p.iread > 0) t = 1.;
  double scal = (10. - 15. * t + 6. * t * t) * cube(t);
  double rscal = 1. - scal;
  for (int kk = ks; kk <= ke; ++kk)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, kk);
      double ajacinv = 1. / v->qj[i];
      v->q1[i] = v->q1[i] * scal + rscal * c->rinf * ajacinv;
      v->q2[i] = v->q2[i] * scal + rscal * c->rinf * c->uinf * ajacinv;
      v->q3[i] = v->q3[i] * scal + rscal * c->rinf * c->vinf * ajacinv;
      v->q4[i] = v->q4[i] * scal + rscal * c->einf * ajacinv;
    }
}

/* bcextp, src/bc.f:1962-2026 */
static void bcextp(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir) {
  double foso = 0.0;
  int iadd = idir >= 0 ? 1 : -1, iadir = abs(idir);
  double* qs[4] = {v->q1, v->q2, v->q3, v->q4};
  if (iadir == 1) {
    for (int jc = 1; jc <= je - js + 1; ++jc) {
      int j = (idir == 1) ? je - jc + 1 : js + jc - 1;
      int j1 = j + iadd, j2 = j1 + iadd;
      for (int k = ks; k <= ke; ++k) {
        size_t i = IX(M, j, k), i1 = IX(M, j1, k), i2 = IX(M, j2, k);
        double s1 = v->qj[i1] / v->qj[i], s2 = v->qj[i2] / v->qj[i];
        for (int n = 0; n < 4; ++n) qs[n][i] = (1. + foso) * qs[n][i1] * s1 - foso * qs[n][i2] * s2;
      }
    }
  } else if (iadir == 2) {
    for (int kc = 1; kc <= ke - ks + 1; ++kc) {
      int k = (idir == 2) ? ke - kc + 1 : ks + kc - 1;
      int k1 = k + iadd, k2 = k1 + iadd;
      for (int j = js; j <= je; ++j) {
        size_t i = IX(M, j, k), i1 = IX(M, j, k1), i2 = IX(M, j, k2);
        double s1 = v->qj[i1] / v->qj[i], s2 = v->qj[i2] / v->qj[i];
        for (int n = 0; n < 4; ++n) qs[n][i] = (1. + foso) * qs[n][i1] * s1 - foso * qs[n][i2] * s2;
      }
    }
  }
}

/* bcsym, src/bc.f:2096-2157 */
static void bcsym(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir) {
  int iadir = abs(idir);
  if (iadir == 1) {
    for (int jc = 1; jc <= je - js + 1; ++jc) {
      int j, j1;
      if (idir == 1) { j = je - jc + 1; j1 = je + jc; } else { j = js + jc - 1; j1 = js - jc; }
      for (int k = ks; k <= ke; ++k) {
        size_t i = IX(M, j, k), i1 = IX(M, j1, k);
        double sc = v->qj[i1] / v->qj[i];
        v->q1[i] = v->q1[i1] * sc; v->q2[i] = -v->q2[i1] * sc;
        v->q3[i] = v->q3[i1] * sc; v->q4[i] = v->q4[i1] * sc;
      }
    }
  } else if (iadir == 2) {
    for (int kc = 1; kc <= ke - ks + 1; ++kc) {
      int k, k1;
      if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
      for (int j = js; j <= je; ++j) {
        size_t i = IX(M, j, k), i1 = IX(M, j, k1);
        double sc = v->qj[i1] / v->qj[i];
        v->q1[i] = v->q1[i1] * sc; v->q2[i] = v->q2[i1] * sc;
        v->q3[i] = -v->q3[i1] * sc; v->q4[i] = v->q4[i1] * sc;
      }
    }
  }
}

/* periodic source index, src/bc.f:192-262 */
static int periodic_src(int idir, int js, int je, int ks, int ke, int jd, int kd, int jk) {
  if (idir == 1) { int jj = je - js + 1; return jd - 2 * jj + (jk - js + 1); }
  if (idir == -1) { int jj = je - js + 1; return 1 + 2 * jj - (je - jk + 1); }
  if (idir == 2) { int kk = ke - ks + 1; return kd - 2 * kk + (jk - ks + 1); }
  int kk = ke - ks + 1;
  return 1 + 2 * kk - (ke - jk + 1);
}

/* bcperiodic, src/bc.f:180-267 */
static void bcperiodic(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir) {
  if (abs(idir) == 1) {
    for (int j = js; j <= je; ++j) {
      int jc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, j);
      for (int k = ks; k <= ke; ++k) {
        size_t i = IX(M, j, k), i1 = IX(M, jc, k);
        double sc = v->qj[i1] / v->qj[i];
        v->q1[i] = v->q1[i1] * sc; v->q2[i] = v->q2[i1] * sc; v->q3[i] = v->q3[i1] * sc; v->q4[i] = v->q4[i1] * sc;
      }
    }
  } else {
    for (int k = ks; k <= ke; ++k) {
      int kc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, k);
      for (int j = js; j <= je; ++j) {
        size_t i = IX(M, j, k), i1 = IX(M, j, kc);
        double sc = v->qj[i1] / v->qj[i];
        v->q1[i] = v->q1[i1] * sc; v->q2[i] = v->q2[i1] * sc; v->q3[i] = v->q3[i1] * sc; v->q4[i] = v->q4[i1] * sc;
      }
    }
  }
}

/* bcwake, src/bc.f:1516-1573 */
static void bcwake(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir) {
  if (abs(idir) != 2) return;
  for (int kc = 1; kc <= ke - ks + 1; ++kc) {
    int k, k1;
    if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
    for (int j = js; j <= je; ++j) {
      int jj = M->jd - j + 1;
      size_t a = IX(M, j, k), b = IX(M, jj, k1), cc = IX(M, jj, k), d = IX(M, j, k1);
      double scale1 = v->qj[b] / v->qj[a];
      double scale2 = v->qj[d] / v->qj[cc];
      v->q1[a] = v->q1[b] * scale1; v->q2[a] = v->q2[b] * scale1;
      v->q3[a] = v->q3[b] * scale1; v->q4[a] = v->q4[b] * scale1;
      v->q1[cc] = v->q1[d] * scale2; v->q2[cc] = v->q2[d] * scale2;
      v->q3[cc] = v->q3[d] * scale2; v->q4[cc] = v->q4[d] * scale2;
    }
  }
}

/* bcout_inf, src/bc.f:360-407 */
static void bcout_inf(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir) {
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double rj = 1. / v->qj[i];
      v->q1[i] = c->rinf * rj;
      v->q2[i] = c->rinf * c->uinf * rj;
      v->q3[i] = c->rinf * c->vinf * rj;
      v->q4[i] = c->pinf / c->gm1 * rj + 0.5 * (sq(v->q2[i]) + sq(v->q3[i])) / v->q1[i];
    }
  (void)idir;
}

/* characteristic far field, bcout_trkl src/bc.f:694-983: one boundary cell
 * i with its first two interior neighbours i1, i2 and outward metric (nx,ny)
 * already normalised as xxn/xyn (or yxn/yyn). */
static void trkl_cell(or_case* c, qview* v, size_t i, size_t i1, size_t i2, double xxn, double xyn) {
  double gm1 = c->gm1, gamma = c->p.gamma, Mp = c->p.Mp;
  double foso = 1.0;
  double qext[4], qbar[4], delq[4], delw[4];
  double* qs[4] = {v->q1, v->q2, v->q3, v->q4};
  for (int n = 0; n < 4; ++n)
    qext[n] = (1. + foso) * qs[n][i1] * v->qj[i1] - foso * qs[n][i2] * v->qj[i2];
  qbar[0] = 0.5 * (qext[0] + c->rinf);
  qbar[1] = 0.5 * (qext[1] + c->rinf * c->uinf);
  qbar[2] = 0.5 * (qext[2] + c->rinf * c->vinf);
  qbar[3] = 0.5 * (qext[3] + c->einf);
  delq[0] = 0.5 * (c->rinf - qext[0]);
  delq[1] = 0.5 * (c->rinf * c->uinf - qext[1]);
  delq[2] = 0.5 * (c->rinf * c->vinf - qext[2]);
  delq[3] = 0.5 * (c->einf - qext[3]);
  double rho = qbar[0];
  double rrho = 1. / rho;
  double uu = qbar[1] * rrho, vv = qbar[2] * rrho, e = qbar[3] * rrho;
  double uv2 = 0.5 * (uu * uu + vv * vv);
  double cjkl = sqrt(c->ggm1 * (e - uv2));
  double c2i = 1. / (cjkl * cjkl);
  double ge = gamma * e - gm1 * uv2;
  double qq = (uu - v->ug[i]) * xxn + (vv - v->vg[i]) * xyn;
  double bt = v->bt[i];
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
  for (int n = 0; n < 4; ++n) qs[n][i] = (qbar[n] - delq[n]) / v->qj[i];
}

/* zeroth-order extrapolation of primitives to further halo layers
 * (src/bc.f:810-835); empty with one halo ring. */
static void extrap_prim(or_case* c, qview* v, size_t i, size_t i1, size_t i2) {
  double gm1 = c->gm1, foso = 0.0;
  double rho1 = v->q1[i1] * v->qj[i1], rho2 = v->q1[i2] * v->qj[i2];
  double u1 = v->q2[i1] / v->q1[i1], u2 = v->q2[i2] / v->q1[i2];
  double v1 = v->q3[i1] / v->q1[i1], v2 = v->q3[i2] / v->q1[i2];
  double p1 = gm1 * (v->q4[i1] - 0.5 * (sq(v->q2[i1]) + sq(v->q3[i1])) / v->q1[i1]) * v->qj[i1];
  double p2 = gm1 * (v->q4[i2] - 0.5 * (sq(v->q2[i2]) + sq(v->q3[i2])) / v->q1[i2]) * v->qj[i2];
  v->q1[i] = ((1. + foso) * rho1 - foso * rho2) / v->qj[i];
  v->q2[i] = ((1. + foso) * u1 - foso * u2) * v->q1[i];
  v->q3[i] = ((1. + foso) * v1 - foso * v2) * v->q1[i];
  double press = (1. + foso) * p1 - foso * p2;
  v->q4[i] = press / (gm1 * v->qj[i]) + 0.5 * (sq(v->q2[i]) + sq(v->q3[i])) / v->q1[i];
}

static void bcout_trkl(or_case* c, or_mesh* M, qview* v, int js, int je, int ks, int ke, int idir) {
  int iadd = idir >= 0 ? 1 : -1, iadir = abs(idir);
  if (iadir == 1) {
    int j = (idir == 1) ? je : js;
    int j1 = j + iadd, j2 = j1 + iadd;
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k);
      double snorm = -1. * iadd / sqrt(sq(v->xx[i]) + sq(v->xy[i]));
      trkl_cell(c, v, i, IX(M, j1, k), IX(M, j2, k), v->xx[i] * snorm, v->xy[i] * snorm);
    }
    for (int jc = 1; jc <= je - js; ++jc) {
      int jj = (idir == 1) ? je - jc : js + jc;
      for (int k = ks; k <= ke; ++k)
        extrap_prim(c, v, IX(M, jj, k), IX(M, jj + iadd, k), IX(M, jj + 2 * iadd, k));
    }
  } else if (iadir == 2) {
    int k = (idir == 2) ? ke : ks;
    int k1 = k + iadd, k2 = k1 + iadd;
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double snorm = -1. * iadd / sqrt(sq(v->yx[i]) + sq(v->yy[i]));
      trkl_cell(c, v, i, IX(M, j, k1), IX(M, j, k2), v->yx[i] * snorm, v->yy[i] * snorm);
    }
    for (int kc = 1; kc <= ke - ks; ++kc) {
      int kk = (idir == 2) ? ke - kc : ks + kc;
      for (int j = js; j <= je; ++j)
        extrap_prim(c, v, IX(M, j, kk), IX(M, j, kk + iadd), IX(M, j, kk + 2 * iadd));
    }
  }
}

/* bc dispatch, src/bc.f:2-85 */
static void bc_mesh(or_case* c, or_mesh* M, int nsp) {
  qview v = view(c, M, nsp);
  for (int ib = 0; ib < M->d.nbc; ++ib) {
    const ot2d_bcseg* b = &M->d.bc[ib];
    int js, je, ks, ke;
    bc_range(M, b, &js, &je, &ks, &ke);
    int idir = b->ibdir;
    switch (b->ibtyp) {
      case 3: bcwall(c, M, &v, js, je, ks, ke, idir, 1); break;
      case 4: bcwall(c, M, &v, js, je, ks, ke, idir, 0); break;
      case 5: bcwall(c, M, &v, js, je, ks, ke, idir, c->p.invisc); break;
      case 10: bcextp(c, M, &v, js, je, ks, ke, idir); break;
      case 11: bcsym(c, M, &v, js, je, ks, ke, idir); break;
      case 22: bcperiodic(c, M, &v, js, je, ks, ke, idir); break;
      case 51: bcwake(c, M, &v, js, je, ks, ke, idir); break;
      case 46: bcout_inf(c, M, &v, js, je, ks, ke, idir); break;
      case 47: bcout_trkl(c, M, &v, js, je, ks, ke, idir); break;
      default: break;
    }
  }
}

int32_t or_bc(or_case* c) {
  for (int im = 0; im < c->nmesh; ++im)
    for (int nsp = 0; nsp < c->nspec; ++nsp) bc_mesh(c, &c->m[im], nsp);
  return OT2D_OK;
}

/* -------------------------------------------------- reconstruction (a2) */
/* muscld_new, src/eulerrhs.f:821-912 (ibmin = ibmax = 1) */
static void muscld_new(double* f, double* ql, double* qr, int mdim, int is, int imax, int im,
                       double th, double qt, double eps, const int* iba, double* f2) {
  if (qt == 0.0) {
    for (int n = 0; n < 4; ++n)
      for (int i = is; i <= imax; ++i) { ql[n * mdim + i] = f[n * mdim + i]; qr[n * mdim + i] = f[n * mdim + i]; }
    return;
  }
  double thm = 1.0 - th, thp = 1.0 + th;
  for (int n = 0; n < 4; ++n) {
    double* fn = f + n * mdim;
    double* l = ql + n * mdim;
    double* r = qr + n * mdim;
    for (int i = is; i <= im; ++i) f2[i] = fn[i + 1] - fn[i];
    for (int i = is + 1; i <= im; ++i) {
      double f2i = f2[i], f2i1 = f2[i - 1];
      double a1 = 3.0 * (f2i * f2i1 + eps);
      double a2 = 2.0 * sq(f2i - f2i1) + a1;
      double f3i = a1 / a2;
      double f3qt = qt * f3i * (double)iba[i] * (double)iba[i + 1] * (double)iba[i - 1];
      l[i] = fn[i] + f3qt * (thm * f2i1 + thp * f2i);
      r[i] = fn[i] - f3qt * (thp * f2i1 + thm * f2i);
    }
    r[is] = fn[is]; l[is] = fn[is];
    l[imax] = fn[imax]; r[imax] = fn[imax];
  }
}

/* ------------------------------------------------------------ WENO5 (a3) */
/* minmod of src/eulerrhs.f:1036 (ammd) */
static double ammd(double fl, double fr) {
  return 0.5 * (copysign(1.0, fl) + copysign(1.0, fr)) * MIN(fabs(fl), fabs(fr));
}

/* weno5, src/eulerrhs.f:1192-1228: Jiang-Shu weights, epsw = 1e-6 */
static double weno5(double a, double b, double c, double d, double e) {
  double b1 = 13. / 12., b2 = 1. / 6., epsw = 1.e-6;
  double djm1 = a - 2. * b + c, ejm1 = a - 4. * b + 3. * c;
  double dj = b - 2. * c + d, ej = b - d;
  double djp1 = c - 2. * d + e, ejp1 = 3. * c - 4. * d + e;
  double dis0 = b1 * djm1 * djm1 + 0.25 * ejm1 * ejm1 + epsw;
  double dis1 = b1 * dj * dj + 0.25 * ej * ej + epsw;
  double dis2 = b1 * djp1 * djp1 + 0.25 * ejp1 * ejp1 + epsw;
  double q30 = 2. * a - 7. * b + 11. * c;
  double q31 = -b + 5. * c + 2. * d;
  double q32 = 2. * c + 5. * d - e;
  double d01 = dis0 / dis1, d02 = dis0 / dis2;
  double a1ba0 = 6. * d01, a2ba0 = 3. * d02;
  double w0 = 1. / (1. + a1ba0 + a2ba0);
  double w1 = a1ba0 * w0;
  double w2 = 1. - w0 - w1;
  return b2 * (w0 * q30 + w1 * q31 + w2 * q32);
}

/* SONIC-A slope at i from limited curvatures (src/eulerrhs.f:1125-1140) */
static double sonic_slope(const double* f1, const double* f2m, int i) {
  double at = f1[i] - 0.5 * f2m[i];
  double ati1 = f1[i - 1] + 0.5 * f2m[i - 1];
  double s1 = ammd(f1[i], f1[i - 1]);
  double t1 = ammd(at, ati1);
  return copysign(1.0, t1) * MIN(0.5 * fabs(at + ati1), MAX(2. * fabs(s1), fabs(t1)));
}

/* weno, src/eulerrhs.f:1011-1189, as rhsup calls it (ibmin = ibmax = 1:
   the fmin/fmax boundary branches are never taken).  Lines of cells
   is..ie; ql(ie) and qr(is) are not used by the caller. */
static void weno_line(const double* f, double* ql, double* qr, int mdim, int is, int ie, double qt,
                      const int* iba, double* wk) {
  if (qt == 0.0) {
    for (int n = 0; n < 4; ++n)
      for (int i = is; i <= ie; ++i) { ql[n * mdim + i] = f[n * mdim + i]; qr[n * mdim + i] = f[n * mdim + i]; }
    return;
  }
  double* f1 = wk;
  double* f2 = wk + mdim;
  double* f2m = wk + 2 * mdim;
  for (int n = 0; n < 4; ++n) {
    const double* f0 = f + n * mdim;
    double* l = ql + n * mdim;
    double* r = qr + n * mdim;
    for (int i = is; i <= ie - 1; ++i) f1[i] = f0[i + 1] - f0[i];
    for (int i = is + 1; i <= ie - 1; ++i) f2[i] = f1[i] - f1[i - 1];
    f2[is] = 2. * f2[is + 1] - f2[is + 2];
    f2[ie] = 2. * f2[ie - 1] - f2[ie - 2];
    for (int i = is; i <= ie - 1; ++i) f2m[i] = ammd(f2[i], f2[i + 1]);
    /* first order-like ends (src/eulerrhs.f:1089-1091, 1104-1106) */
    double at = f1[is] - 0.5 * f2m[is];
    l[is] = f0[is] + 0.5 * ammd(at, 2. * f1[is]);
    double at1 = f1[ie - 1] + 0.5 * f2m[ie - 1];
    r[ie] = f0[ie] - 0.5 * ammd(at1, 2. * f1[ie - 1]);
    /* SONIC-A at is+1 and ie-1 (src/eulerrhs.f:1121-1141) */
    for (int i = is + 1; i <= ie - 1; i += ie - is - 2) {
      double sl = sonic_slope(f1, f2m, i);
      l[i] = f0[i] + 0.5 * sl;
      r[i] = f0[i] - 0.5 * sl;
    }
    /* WENO5 inside, first order next to blanked cells (1145-1153) */
    for (int i = is + 2; i <= ie - 2; ++i) {
      l[i] = weno5(f0[i - 2], f0[i - 1], f0[i], f0[i + 1], f0[i + 2]);
      r[i] = weno5(f0[i + 2], f0[i + 1], f0[i], f0[i - 1], f0[i - 2]);
      double blankv = (double)(iba[i] * iba[i - 1] * iba[i + 1] * iba[i - 2] * iba[i + 2]);
      l[i] = l[i] * blankv + (1. - blankv) * f0[i];
      r[i] = r[i] * blankv + (1. - blankv) * f0[i];
    }
    /* SONIC-A again for ql(is+2) and qr(ie-1) (1157-1187) */
    l[is + 2] = f0[is + 2] + 0.5 * sonic_slope(f1, f2m, is + 2);
    r[ie - 1] = f0[ie - 1] - 0.5 * sonic_slope(f1, f2m, ie - 1);
  }
}

/* ---------------------------------------------------------- Roe (a4) */
/* roeflx, src/eulerrhs.f:283-446 */
static void roeflx(or_case* c, double* f, const double* ql, const double* qr, int mdim,
                   const double* xa, const double* ya, const double* tj, int is, int ie) {
  double gm1 = c->gm1;
  double eps = 1.e-6;
  for (int i = is; i <= ie; ++i) {
    int i1 = i + 1;
    double rlft = ql[i], ulft = ql[mdim + i], vlft = ql[2 * mdim + i], plft = ql[3 * mdim + i];
    double rlfti = 1.0 / rlft;
    double rulft = rlft * ulft, rvlft = rlft * vlft;
    double uvl = 0.5 * (ulft * ulft + vlft * vlft);
    double elft = plft / gm1 + rlft * uvl;
    double hlft = (elft + plft) * rlfti;
    double clft = sqrt(gm1 * (hlft - uvl));
    double rrht = qr[i1], urht = qr[mdim + i1], vrht = qr[2 * mdim + i1], prht = qr[3 * mdim + i1];
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
    double ri1 = xa[i], ri2 = ya[i], ri3 = tj[i];
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
    double dauus = MAX(dauu, 0.0);
    if (auu <= 0.5 * dauus) auu = sq(auu) / dauu + 0.25 * dauu;
    double daupc = 4. * (upcrht - upclft) + eps;
    double daupcs = MAX(daupc, 0.0);
    if (aupc <= 0.5 * daupcs) aupc = sq(aupc) / daupc + 0.25 * daupc;
    double daumc = 4. * (umcrht - umclft) + eps;
    double daumcs = MAX(daumc, 0.0);
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
    f[i] = aj * (rlft * uulft + rrht * uurht - aq1) - fssub * c->rinf;
    f[mdim + i] = aj * (rulft * uulft + rurht * uurht + r1 * plar - aq2) - fssub * c->rinf * c->uinf;
    f[2 * mdim + i] = aj * (rvlft * uulft + rvrht * uurht + r2 * plar - aq3) - fssub * c->rinf * c->vinf;
    f[3 * mdim + i] = aj * (eplft * uulft + eprht * uurht - r3 * plar - aq4) - fssub * c->einf;
  }
}

/* --------------------------------------------------- inviscid rhs (a1) */
/* rhsup, src/eulerrhs.f:2-280 (nei = 1) */
static int rhsup_mesh(or_case* c, or_mesh* M, int nsp) {
  int ilima = abs(c->p.ilim);
  if (!(ilima == 1 || c->p.ilim == 4)) return OT2D_ERR_UNSUPPORTED;
  int irhs = abs(c->p.irhsy);
  double th = (double)(irhs - 2) / (double)irhs;
  double qt = 0.25;
  if (irhs == 1) qt = 0.0;
  double epsj = cube(10. / (double)M->jmax);
  double epsk = cube(10. / (double)M->kmax);
  int mdim = c->mdim + 2;
  double* f = (double*)calloc((size_t)mdim * 4, sizeof(double));
  double* ql = (double*)calloc((size_t)mdim * 4, sizeof(double));
  double* qr = (double*)calloc((size_t)mdim * 4, sizeof(double));
  double* f2 = (double*)calloc((size_t)mdim * 3, sizeof(double));
  double* xa = (double*)calloc((size_t)mdim, sizeof(double));
  double* ya = (double*)calloc((size_t)mdim, sizeof(double));
  double* tj = (double*)calloc((size_t)mdim, sizeof(double));
  int* iba = (int*)calloc((size_t)mdim, sizeof(int));
  double* ppp = (double*)calloc(M->n2, sizeof(double));
  qview v = view(c, M, nsp);
  const int32_t* ib = M->iblank + (size_t)nsp * M->n2;
  const double* xv = VA(M, xv, nsp);
  const double* yv = VA(M, yv, nsp);
  const double* ugv = VA(M, ugv, nsp);
  const double* vgv = VA(M, vgv, nsp);
  double *s1 = SA(M, nsp, 1), *s2 = SA(M, nsp, 2), *s3 = SA(M, nsp, 3), *s4 = SA(M, nsp, 4);
  double* ss[4] = {s1, s2, s3, s4};
  for (size_t i = 0; i < M->n2; ++i)
    ppp[i] = c->gm1 * (v.q4[i] - 0.5 * (sq(v.q2[i]) + sq(v.q3[i])) / v.q1[i]);
  /* xi fluxes along k = const lines */
  for (int k = M->kbeg; k <= M->kend; ++k) {
    int kv = k;
    for (int j = 1; j <= M->jd; ++j) {
      size_t i = IX(M, j, k);
      double rhoi = 1.0 / v.q1[i];
      f[j] = v.q1[i] * v.qj[i];
      f[mdim + j] = v.q2[i] * rhoi;
      f[2 * mdim + j] = v.q3[i] * rhoi;
      f[3 * mdim + j] = ppp[i] * v.qj[i];
      iba[j] = abs(ib[i]);
    }
    if (ilima == 1) muscld_new(f, ql, qr, mdim, 1, M->jd, M->jd - 1, th, qt, epsj, iba, f2);
    if (c->p.ilim == 4) weno_line(f, ql, qr, mdim, 1, M->jd, qt, iba, f2);
    for (int j = M->jbeg - 1; j <= M->jend; ++j) {
      int jv = j;
      xa[j] = (yv[VX(M, jv, kv)] - yv[VX(M, jv, kv - 1)]);
      ya[j] = -(xv[VX(M, jv, kv)] - xv[VX(M, jv, kv - 1)]);
    }
    if (c->timespectral) {
      for (int j = M->jbeg - 1; j <= M->jend; ++j) {
        int jv = j;
        tj[j] = -0.5 * (ugv[VX(M, jv, kv)] + ugv[VX(M, jv, kv - 1)]) * xa[j]
                - 0.5 * (vgv[VX(M, jv, kv)] + vgv[VX(M, jv, kv - 1)]) * ya[j];
      }
    } else {
      for (int j = M->jbeg - 1; j <= M->jend; ++j) tj[j] = 0.0;
    }
    roeflx(c, f, ql, qr, mdim, xa, ya, tj, M->jbeg - 1, M->jend);
    for (int j = M->jbeg; j <= M->jend; ++j) {
      size_t i = IX(M, j, k);
      for (int n = 0; n < 4; ++n) ss[n][i] = ss[n][i] - (f[n * mdim + j] - f[n * mdim + j - 1]);
    }
  }
  /* eta fluxes along j = const lines */
  for (int j = M->jbeg; j <= M->jend; ++j) {
    int jv = j;
    for (int k = 1; k <= M->kd; ++k) {
      size_t i = IX(M, j, k);
      double rhoi = 1.0 / v.q1[i];
      f[k] = v.q1[i] * v.qj[i];
      f[mdim + k] = v.q2[i] * rhoi;
      f[2 * mdim + k] = v.q3[i] * rhoi;
      f[3 * mdim + k] = ppp[i] * v.qj[i];
      iba[k] = abs(ib[i]);
    }
    if (ilima == 1) muscld_new(f, ql, qr, mdim, 1, M->kd, M->kd - 1, th, qt, epsk, iba, f2);
    if (c->p.ilim == 4) weno_line(f, ql, qr, mdim, 1, M->kd, qt, iba, f2);
    for (int k = M->kbeg - 1; k <= M->kend; ++k) {
      int kv = k;
      xa[k] = -(yv[VX(M, jv, kv)] - yv[VX(M, jv - 1, kv)]);
      ya[k] = (xv[VX(M, jv, kv)] - xv[VX(M, jv - 1, kv)]);
    }
    if (c->timespectral) {
      for (int k = M->kbeg - 1; k <= M->kend; ++k) {
        int kv = k;
        tj[k] = -0.5 * (ugv[VX(M, jv, kv)] + ugv[VX(M, jv - 1, kv)]) * xa[k]
                - 0.5 * (vgv[VX(M, jv, kv)] + vgv[VX(M, jv - 1, kv)]) * ya[k];
      }
    } else {
      for (int k = M->kbeg - 1; k <= M->kend; ++k) tj[k] = 0.;
    }
    roeflx(c, f, ql, qr, mdim, xa, ya, tj, M->kbeg - 1, M->kend);
    for (int k = M->kbeg; k <= M->kend; ++k) {
      size_t i = IX(M, j, k);
      for (int n = 0; n < 4; ++n) ss[n][i] = ss[n][i] - (f[n * mdim + k] - f[n * mdim + k - 1]);
    }
  }
  free(f); free(ql); free(qr); free(f2); free(xa); free(ya); free(tj); free(iba); free(ppp);
  return OT2D_OK;
}

/* timespectralrhs, src/sources.f:2-27 */
static void timespectralrhs(or_case* c, or_mesh* M, int nsp) {
  int N = c->nspec;
  for (int n = 1; n <= c->nv; ++n) {
    double* s = SA(M, nsp, n);
    for (int nsp1 = 0; nsp1 < N; ++nsp1) {
      double dsv = c->Ds[nsp + (size_t)nsp1 * N];
      const double* qn = QA(M, nsp1, n);
      for (int k = M->kbeg; k <= M->kend; ++k)
        for (int j = M->jbeg; j <= M->jend; ++j) {
          size_t i = IX(M, j, k);
          s[i] = s[i] - M->rotf * dsv * qn[i];
        }
    }
  }
}

/* ------------------------------------------------------- viscous (a5) */
/* visrhs, src/visrhs.f:2-287 (nhalo = 1) */
static void visrhs_mesh(or_case* c, or_mesh* M, int nsp) {
  int jd = M->jd, kd = M->kd;
  double gamma = c->p.gamma, pr = c->p.pr;
  double gkpr = gamma / pr, prtr = pr / 0.9, dre = .5 / c->rey;
  double c2b = 198.6 / c->p.tinf, c2bp = c2b + 1.;
  double ggm1 = c->ggm1;
  double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
  double* qj = QA(M, nsp, c->nq);
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  double* turmu = GA(M, turmu, nsp);
  double *s2 = SA(M, nsp, 2), *s3 = SA(M, nsp, 3), *s4 = SA(M, nsp, 4);
  int md = c->mdim + 2;
  double *f2 = calloc(md, 8), *f3 = calloc(md, 8), *f4 = calloc(md, 8);
  double *f2v = calloc(md, 8), *f3v = calloc(md, 8), *f4v = calloc(md, 8);
  double *u = calloc(md, 8), *v = calloc(md, 8), *e = calloc(md, 8);
  double *a1v = calloc(md, 8), *a2v = calloc(md, 8), *a4v = calloc(md, 8), *a5v = calloc(md, 8);
  double *d1v = calloc(md, 8), *d2v = calloc(md, 8), *d4v = calloc(md, 8), *d5v = calloc(md, 8), *d7v = calloc(md, 8);
#define I2(j, k) IX(M, j, k)
  for (int j = M->jbeg; j <= M->jend; ++j) {
    for (int k = 1; k <= kd; ++k) {
      size_t i = I2(j, k), im = I2(j - 1, k), ip = I2(j + 1, k);
      double ra = 1. / q1[i];
      u[k] = q2[i] * ra;
      v[k] = q3[i] * ra;
      e[k] = q4[i] * ra - .5 * (sq(u[k]) + sq(v[k]));
      double tt = ggm1 * e[k];
      double vmue = c2bp * tt * sqrt(tt) / (c2b + tt);
      double turm = turmu[i];
      double vnu = vmue + turm;
      double gkap = vmue + prtr * turm;
      double rj = 1. / qj[i], rjm = 1. / qj[im], rjp = 1. / qj[ip];
      double b4 = (sq(yx[i]) + sq(yy[i])) * rj;
      a1v[k] = (b4 + sq(yx[i]) / 3. * rj) * vnu * dre;   /* b1v */
      a2v[k] = (b4 + sq(yy[i]) / 3. * rj) * vnu * dre;   /* b2v */
      a5v[k] = (yx[i] * yy[i] / 3. * rj) * vnu * dre;    /* b5v */
      a4v[k] = b4 * gkpr * gkap * dre;                   /* b4v */
      double t;
      t = (4. / 3. * xx[im] * yx[im] + xy[im] * yy[im]) * rjm;
      t = t + (4. / 3. * xx[ip] * yx[ip] + xy[ip] * yy[ip]) * rjp;
      d1v[k] = vnu * dre * t;
      t = (4. / 3. * xy[im] * yy[im] + xx[im] * yx[im]) * rjm;
      t = t + (4. / 3. * xy[ip] * yy[ip] + xx[ip] * yx[ip]) * rjp;
      d2v[k] = vnu * dre * t;
      t = (xy[im] * yy[im] + xx[im] * yx[im]) * rjm;
      t = t + (xy[ip] * yy[ip] + xx[ip] * yx[ip]) * rjp;
      d4v[k] = t * gkpr * gkap * dre;
      t = (-2. / 3. * xy[im] * yx[im] + xx[im] * yy[im]) * rjm;
      t = t + (-2. / 3. * xy[ip] * yx[ip] + xx[ip] * yy[ip]) * rjp;
      d5v[k] = vnu * dre * t;
      t = (-2. / 3. * xx[im] * yy[im] + xy[im] * yx[im]) * rjm;
      t = t + (-2. / 3. * xx[ip] * yy[ip] + xy[ip] * yx[ip]) * rjp;
      d7v[k] = vnu * dre * t;
    }
    for (int k = 1; k <= kd - 1; ++k) {
      int k1 = k + 1;
      double b1 = a1v[k1] + a1v[k], b2 = a2v[k1] + a2v[k], b5 = a5v[k1] + a5v[k], b4 = a4v[k1] + a4v[k];
      double d1 = d1v[k], d2 = d2v[k], d4 = d4v[k], d5 = d5v[k], d7 = d7v[k];
      size_t ip = I2(j + 1, k), im = I2(j - 1, k);
      double ujp1 = q2[ip] / q1[ip], ujm1 = q2[im] / q1[im];
      double duj = 0.5 * (ujp1 - ujm1);
      double vjp1 = q3[ip] / q1[ip], vjm1 = q3[im] / q1[im];
      double dvj = 0.5 * (vjp1 - vjm1);
      double ejp1 = q4[ip] / q1[ip] - .5 * (sq(ujp1) + sq(vjp1));
      double ejm1 = q4[im] / q1[im] - .5 * (sq(ujm1) + sq(vjm1));
      double dej = 0.5 * (ejp1 - ejm1);
      double t4 = u[k1] * a1v[k1] + u[k] * a1v[k];
      double t5 = v[k1] * a2v[k1] + v[k] * a2v[k];
      double t6 = u[k1] * a5v[k1] + u[k] * a5v[k];
      double t7 = v[k1] * a5v[k1] + v[k] * a5v[k];
      double t8 = a4v[k1] + a4v[k];
      double du = u[k1] - u[k], dv = v[k1] - v[k], dei = e[k1] - e[k];
      f2[k] = b1 * du + b5 * dv;
      f2v[k] = d1 * duj + d5 * dvj;
      f3[k] = b5 * du + b2 * dv;
      f3v[k] = d7 * duj + d2 * dvj;
      f4[k] = (t4 + t7) * du + (t5 + t6) * dv + t8 * dei;
      f4v[k] = 0.5 * d1 * 0.5 * (sq(ujp1) - sq(ujm1));
      f4v[k] = f4v[k] + 0.5 * d2 * 0.5 * (sq(vjp1) - sq(vjm1));
      f4v[k] = f4v[k] + d5 * 0.5 * (ujp1 + ujm1) * dvj;
      f4v[k] = f4v[k] + d7 * 0.5 * (vjp1 + vjm1) * duj;
      f4v[k] = f4v[k] + d4 * dej;
    }
    int ke = M->kend - 1;
    int k = M->kend;
    size_t i = I2(j, k);
    s2[i] = s2[i] + (f2[k] - f2[k - 1]) + 0.5 * (f2v[k - 2] - 4. * f2v[k - 1] + 3 * f2v[k]);
    s3[i] = s3[i] + (f3[k] - f3[k - 1]) + 0.5 * (f3v[k - 2] - 4. * f3v[k - 1] + 3 * f3v[k]);
    s4[i] = s4[i] + (f4[k] - f4[k - 1]) + 0.5 * (f4v[k - 2] - 4. * f4v[k - 1] + 3. * f4v[k]);
    for (k = M->kbeg; k <= ke; ++k) {
      i = I2(j, k);
      s2[i] = s2[i] + (f2[k] - f2[k - 1]) + 0.5 * (f2v[k + 1] - f2v[k - 1]);
      s3[i] = s3[i] + (f3[k] - f3[k - 1]) + 0.5 * (f3v[k + 1] - f3v[k - 1]);
      s4[i] = s4[i] + (f4[k] - f4[k - 1]) + 0.5 * (f4v[k + 1] - f4v[k - 1]);
    }
  }
  for (int k = M->kbeg; k <= M->kend; ++k) {
    for (int j = 1; j <= jd; ++j) {
      size_t i = I2(j, k), km = I2(j, k - 1), kp = I2(j, k + 1);
      double ra = 1. / q1[i];
      u[j] = q2[i] * ra;
      v[j] = q3[i] * ra;
      e[j] = q4[i] * ra - .5 * (sq(u[j]) + sq(v[j]));
      double tt = ggm1 * e[j];
      double vmue = c2bp * tt * sqrt(tt) / (c2b + tt);
      double turm = turmu[i];
      double vnu = vmue + turm;
      double gkap = vmue + prtr * turm;
      double rj = 1. / qj[i], rkm = 1. / qj[km], rkp = 1. / qj[kp];
      double a4 = (sq(xx[i]) + sq(xy[i])) * rj;
      a1v[j] = (a4 + sq(xx[i]) / 3. * rj) * vnu * dre;
      a2v[j] = (a4 + sq(xy[i]) / 3. * rj) * vnu * dre;
      a5v[j] = (xx[i] * xy[i] / 3. * rj) * vnu * dre;
      a4v[j] = a4 * gkpr * gkap * dre;
      double t;
      t = (4. / 3. * xx[km] * yx[km] + xy[km] * yy[km]) * rkm;
      t = t + (4. / 3. * xx[kp] * yx[kp] + xy[kp] * yy[kp]) * rkp;
      d1v[j] = vnu * dre * t;
      t = (4. / 3. * xy[km] * yy[km] + xx[km] * yx[km]) * rkm;
      t = t + (4. / 3. * xy[kp] * yy[kp] + xx[kp] * yx[kp]) * rkp;
      d2v[j] = vnu * dre * t;
      t = (xy[km] * yy[km] + xx[km] * yx[km]) * rkm;
      t = t + (xy[kp] * yy[kp] + xx[kp] * yx[kp]) * rkp;
      d4v[j] = t * gkpr * gkap * dre;
      t = (-2. / 3. * xy[km] * yx[km] + xx[km] * yy[km]) * rkm;
      t = t + (-2. / 3. * xy[kp] * yx[kp] + xx[kp] * yy[kp]) * rkp;
      d5v[j] = vnu * dre * t;
      t = (-2. / 3. * xx[km] * yy[km] + xy[km] * yx[km]) * rkm;
      t = t + (-2. / 3. * xx[kp] * yy[kp] + xy[kp] * yx[kp]) * rkp;
      d7v[j] = vnu * dre * t;
    }
    for (int j = 1; j <= jd - 1; ++j) {
      int j1 = j + 1;
      double a1 = a1v[j1] + a1v[j], a2 = a2v[j1] + a2v[j], a5 = a5v[j1] + a5v[j], a4 = a4v[j1] + a4v[j];
      double d1 = d1v[j], d2 = d2v[j], d4 = d4v[j], d5 = d5v[j], d7 = d7v[j];
      size_t kp = I2(j, k + 1), km = I2(j, k - 1);
      double ukp1 = q2[kp] / q1[kp], ukm1 = q2[km] / q1[km];
      double duk = 0.5 * (ukp1 - ukm1);
      double vkp1 = q3[kp] / q1[kp], vkm1 = q3[km] / q1[km];
      double dvk = 0.5 * (vkp1 - vkm1);
      double ekp1 = q4[kp] / q1[kp] - .5 * (sq(ukp1) + sq(vkp1));
      double ekm1 = q4[km] / q1[km] - .5 * (sq(ukm1) + sq(vkm1));
      double dek = 0.5 * (ekp1 - ekm1);
      double t4 = u[j1] * a1v[j1] + u[j] * a1v[j];
      double t5 = v[j1] * a2v[j1] + v[j] * a2v[j];
      double t6 = u[j1] * a5v[j1] + u[j] * a5v[j];
      double t7 = v[j1] * a5v[j1] + v[j] * a5v[j];
      double t8 = a4v[j1] + a4v[j];
      double du = u[j1] - u[j], dv = v[j1] - v[j], dei = e[j1] - e[j];
      f2[j] = a1 * du + a5 * dv;
      f2v[j] = d1 * duk + d7 * dvk;
      f3[j] = a5 * du + a2 * dv;
      f3v[j] = d5 * duk + d2 * dvk;
      f4[j] = (t4 + t7) * du + (t5 + t6) * dv + t8 * dei;
      f4v[j] = 0.5 * d1 * 0.5 * (sq(ukp1) - sq(ukm1));
      f4v[j] = f4v[j] + 0.5 * d2 * 0.5 * (sq(vkp1) - sq(vkm1));
      f4v[j] = f4v[j] + d5 * 0.5 * (vkp1 + vkm1) * duk;
      f4v[j] = f4v[j] + d7 * 0.5 * (ukp1 + ukm1) * dvk;
      f4v[j] = f4v[j] + d4 * dek;
    }
    int je = M->jend - 1;
    int j = M->jend;
    size_t i = I2(j, k);
    s2[i] = s2[i] + (f2[j] - f2[j - 1]) + 0.5 * (f2v[j - 2] - 4. * f2v[j - 1] + 3 * f2v[j]);
    s3[i] = s3[i] + (f3[j] - f3[j - 1]) + 0.5 * (f3v[j - 2] - 4. * f3v[j - 1] + 3 * f3v[j]);
    s4[i] = s4[i] + (f4[j] - f4[j - 1]) + 0.5 * (f4v[j - 2] - 4. * f4v[j - 1] + 3 * f4v[j]);
    for (j = M->jbeg; j <= je; ++j) {
      i = I2(j, k);
      s2[i] = s2[i] + (f2[j] - f2[j - 1]) + 0.5 * (f2v[j + 1] - f2v[j - 1]);
      s3[i] = s3[i] + (f3[j] - f3[j - 1]) + 0.5 * (f3v[j + 1] - f3v[j - 1]);
      s4[i] = s4[i] + (f4[j] - f4[j - 1]) + 0.5 * (f4v[j + 1] - f4v[j - 1]);
    }
  }
#undef I2
  free(f2); free(f3); free(f4); free(f2v); free(f3v); free(f4v); free(u); free(v); free(e);
  free(a1v); free(a2v); free(a4v); free(a5v); free(d1v); free(d2v); free(d4v); free(d5v); free(d7v);
}

/* ------------------------------------------------- SA turbulence (a12) */
/* constants of src/sadata.h:2-4 */
static const double SA_cb1 = 0.1355, SA_sigma = 0.6666667, SA_cb2 = 0.622, SA_akt = 0.41,
                    SA_cw1 = 3.2390678, SA_cw2 = 0.3, SA_cw3 = 2., SA_cv1 = 7.1, SA_ct3 = 1.2,
                    SA_ct4 = 0.5;

/* integer powers as the reference build evaluates them: repeated
 * multiplication left to right (x**6 = ((((x*x)*x)*x)*x)*x); real powers
 * (x**(1./6.), x**5.) go to the C library pow */
static double ipow3(double x) { return (x * x) * x; }
static double ipow6(double x) { return ((((x * x) * x) * x) * x) * x; }
static double ipow2(double x) { return x * x; }

/* turbc, src/samodel2d.f:827-1303 */
static void turbc_mesh(or_case* c, or_mesh* M, int nsp, const double* u, const double* v) {
  double* q5 = QA(M, nsp, 5);
  double *ug = GA(M, ug, nsp), *vg = GA(M, vg, nsp);
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  for (int ib = 0; ib < M->d.nbc; ++ib) {
    const ot2d_bcseg* b = &M->d.bc[ib];
    int js, je, ks, ke;
    bc_range(M, b, &js, &je, &ks, &ke);
    int idir = b->ibdir, iadd = idir >= 0 ? 1 : -1, iadir = abs(idir);
    int t = b->ibtyp;
    double foso = 1.0;
    if (t == 3 || t == 4) {          /* turbc_extpt */
      if (iadir == 1) {
        for (int jc = 1; jc <= je - js + 1; ++jc) {
          int j = (idir == 1) ? je - jc + 1 : js + jc - 1;
          int j1 = j + iadd, j2 = j1 + iadd;
          for (int k = ks; k <= ke; ++k)
            q5[IX(M, j, k)] = (1. + foso) * q5[IX(M, j1, k)] - foso * q5[IX(M, j2, k)];
        }
      } else if (iadir == 2) {
        for (int kc = 1; kc <= ke - ks + 1; ++kc) {
          int k = (idir == 2) ? ke - kc + 1 : ks + kc - 1;
          int k1 = k + iadd, k2 = k1 + iadd;
          for (int j = js; j <= je; ++j)
            q5[IX(M, j, k)] = (1. + foso) * q5[IX(M, j, k1)] - foso * q5[IX(M, j, k2)];
        }
      }
    } else if (t == 5) {             /* turbc_wall (one halo layer) */
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd;
        for (int k = ks; k <= ke; ++k) q5[IX(M, j, k)] = -q5[IX(M, j1, k)];
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd;
        for (int j = js; j <= je; ++j) q5[IX(M, j, k)] = -q5[IX(M, j, k1)];
      }
    } else if (t == 10 || t == 11) { /* turbc_sym */
      if (iadir == 1) {
        for (int jc = 1; jc <= je - js + 1; ++jc) {
          int j, j1;
          if (idir == 1) { j = je - jc + 1; j1 = je + jc; } else { j = js + jc - 1; j1 = js - jc; }
          for (int k = ks; k <= ke; ++k) q5[IX(M, j, k)] = q5[IX(M, j1, k)];
        }
      } else if (iadir == 2) {
        for (int kc = 1; kc <= ke - ks + 1; ++kc) {
          int k, k1;
          if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
          for (int j = js; j <= je; ++j) q5[IX(M, j, k)] = q5[IX(M, j, k1)];
        }
      }
    } else if (t == 22) {            /* turbc_periodic */
      if (iadir == 1) {
        for (int j = js; j <= je; ++j) {
          int jc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, j);
          for (int k = ks; k <= ke; ++k) q5[IX(M, j, k)] = q5[IX(M, jc, k)];
        }
      } else {
        for (int k = ks; k <= ke; ++k) {
          int kc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, k);
          for (int j = js; j <= je; ++j) q5[IX(M, j, k)] = q5[IX(M, j, kc)];
        }
      }
    } else if (t == 51 && iadir == 2) { /* turbc_wake */
      for (int kc = 1; kc <= ke - ks + 1; ++kc) {
        int k, k1;
        if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
        for (int j = js; j <= je; ++j) {
          int jj = M->jd - j + 1;
          q5[IX(M, j, k)] = q5[IX(M, jj, k1)];
          q5[IX(M, jj, k)] = q5[IX(M, j, k1)];
        }
      }
    } else if (t == 46) {            /* turbc_inf */
      for (int k = ks; k <= ke; ++k)
        for (int j = js; j <= je; ++j) q5[IX(M, j, k)] = c->p.vnuinf;
    } else if (t == 47) {            /* turbc_out */
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd;
        for (int k = ks; k <= ke; ++k) {
          size_t i = IX(M, j, k), i1 = IX(M, j1, k);
          q5[i] = c->p.vnuinf;
          double uu = (u[i] - ug[i] + u[i1] - ug[i1]) * xx[i];
          uu = uu + (v[i] - vg[i] + v[i1] - vg[i1]) * xy[i];
          uu = uu * iadd;
          if (uu < 0.) q5[i] = q5[i1];
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd;
        for (int j = js; j <= je; ++j) {
          size_t i = IX(M, j, k), i1 = IX(M, j, k1);
          q5[i] = c->p.vnuinf;
          double uu = (u[i] - ug[i] + u[i1] - ug[i1]) * yx[i];
          uu = uu + (v[i] - vg[i] + v[i1] - vg[i1]) * yy[i];
          uu = uu * iadd;
          if (uu < 0.) q5[i] = q5[i1];
        }
      }
    }
  }
}

/* vmu_sa, src/samodel2d.f:2-197, itrans = 0, nturiter = nhuang = 1 */
static void vmu_sa_mesh(or_case* c, or_mesh* M, int nsp) {
  int jd = M->jd, kd = M->kd, js = M->jbeg, je = M->jend, ks = M->kbeg, ke = M->kend;
  size_t n2 = M->n2;
  double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
  double *q5 = QA(M, nsp, 5), *qj = QA(M, nsp, c->nq);
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  double *ug = GA(M, ug, nsp), *vg = GA(M, vg, nsp);
  double *x = GA(M, x, nsp), *y = GA(M, y, nsp);
  const double *xv = VA(M, xv, nsp), *yv = VA(M, yv, nsp);
  double* turmu = GA(M, turmu, nsp);
  double* tscale = GA(M, tscale, nsp);
  const int32_t* ib = M->iblank + (size_t)nsp * n2;
  double* s5 = SA(M, nsp, 5);
  double *sjmat = calloc(n2, 8), *bjmat = calloc(n2, 8);
  double *aj = calloc(n2, 8), *bj = calloc(n2, 8), *cj = calloc(n2, 8);
  double *ak = calloc(n2, 8), *bk = calloc(n2, 8), *ck = calloc(n2, 8);
  double *vmul = calloc(n2, 8), *vort = calloc(n2, 8), *sn = calloc(n2, 8);
  double *u = calloc(n2, 8), *v = calloc(n2, 8), *s1 = calloc(n2, 8), *s2 = calloc(n2, 8);
  double *up = calloc(n2, 8), *um = calloc(n2, 8), *vp = calloc(n2, 8), *vm = calloc(n2, 8);
  double *chp = calloc(n2, 8), *dnuhp = calloc(n2, 8), *dmxh = calloc(n2, 8), *dmyh = calloc(n2, 8);
  double relfac = 1.0, vnulim = 1.0e-20;
  double h = c->h;
  /* lamvis, src/samodel2d.f:200-238 */
  {
    double c2b = 198.6 / c->p.tinf, c2bp = c2b + 1.;
    for (size_t i = 0; i < n2; ++i) {
      double ra = 1. / q1[i];
      double uvel = q2[i] * ra, vvel = q3[i] * ra;
      double ei = q4[i] * ra - 0.5 * (sq(uvel) + sq(vvel));
      double tt = c->ggm1 * ei;
      vmul[i] = c2bp * tt * sqrt(tt) / (c2b + tt);
    }
  }
  for (size_t i = 0; i < n2; ++i) { u[i] = q2[i] / q1[i]; v[i] = q3[i] / q1[i]; sn[i] = 1.e10; }
  turbc_mesh(c, M, nsp, u, v);
  if (M->bodyflag) {
    /* vortic, src/samodel2d.f:296-333 */
    for (int k = ks; k <= ke; ++k)
      for (int j = js; j <= je; ++j) {
        size_t i = IX(M, j, k);
        double ux = (u[IX(M, j + 1, k)] - u[IX(M, j - 1, k)]) * 0.5;
        double vx = (v[IX(M, j + 1, k)] - v[IX(M, j - 1, k)]) * 0.5;
        double uy = (u[IX(M, j, k + 1)] - u[IX(M, j, k - 1)]) * 0.5;
        double vy = (v[IX(M, j, k + 1)] - v[IX(M, j, k - 1)]) * 0.5;
        double tx = xy[i] * ux - xx[i] * vx + yy[i] * uy - yx[i] * vy;
        vort[i] = fabs(tx);
      }
    /* dist, src/samodel2d.f:243-293 */
    for (int j = M->jtail1; j <= M->jtail2; ++j) sn[IX(M, j, 1)] = 0.0;
    for (int j = M->jtail1; j <= M->jtail2; ++j)
      for (int k = 2; k <= kd; ++k) {
        int jv = j - 1;
        size_t i = IX(M, j, k);
        sn[i] = sqrt(sq(x[i] - 0.5 * (xv[VX(M, jv, 1)] + xv[VX(M, jv + 1, 1)])) +
                     sq(y[i] - 0.5 * (yv[VX(M, jv, 1)] + yv[VX(M, jv + 1, 1)])));
      }
    int jv = M->jtail1 - 1;
    for (int j = 1; j <= M->jtail1 - 1; ++j)
      for (int k = 1; k <= kd; ++k) {
        size_t i = IX(M, j, k);
        sn[i] = sqrt(sq(x[i] - xv[VX(M, jv, 1)]) + sq(y[i] - yv[VX(M, jv, 1)]));
      }
    for (int j = M->jtail2 + 1; j <= jd; ++j)
      for (int k = 1; k <= kd; ++k) {
        size_t i = IX(M, j, k);
        sn[i] = sqrt(sq(x[i] - xv[VX(M, jv, 1)]) + sq(y[i] - yv[VX(M, jv, 1)]));
      }
  }
  /* rhslhs: sjmat = s5; convec, diffus, source (src/samodel2d.f:402-824) */
  for (size_t i = 0; i < n2; ++i) sjmat[i] = s5[i];
  /* convec, src/samodel2d.f:453-542 */
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      double uu = xx[i] * (u[i] - ug[i]) + xy[i] * (v[i] - vg[i]);
      double vv = yx[i] * (u[i] - ug[i]) + yy[i] * (v[i] - vg[i]);
      up[i] = 0.5 * (uu + fabs(uu)); um[i] = 0.5 * (uu - fabs(uu));
      vp[i] = 0.5 * (vv + fabs(vv)); vm[i] = 0.5 * (vv - fabs(vv));
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      sjmat[i] = sjmat[i] - up[i] * (q5[i] - q5[IX(M, j - 1, k)]) - um[i] * (q5[IX(M, j + 1, k)] - q5[i]);
      sjmat[i] = sjmat[i] - vp[i] * (q5[i] - q5[IX(M, j, k - 1)]) - vm[i] * (q5[IX(M, j, k + 1)] - q5[i]);
    }
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double fwd = up[IX(M, j + 1, k)] > 1.0e-12 ? 1.0 : 0.0;
      double bck = um[IX(M, j - 1, k)] < -1.0e-12 ? 1.0 : 0.0;
      aj[IX(M, j + 1, k)] = aj[IX(M, j + 1, k)] - fwd * (up[i] + um[i]);
      bj[i] = bj[i] + up[i] - um[i];
      cj[IX(M, j - 1, k)] = cj[IX(M, j - 1, k)] + bck * (up[i] + um[i]);
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      double fwd = vp[IX(M, j, k + 1)] > 1.0e-12 ? 1.0 : 0.0;
      double bck = vm[IX(M, j, k - 1)] < -1.0e-12 ? 1.0 : 0.0;
      ak[IX(M, j, k + 1)] = ak[IX(M, j, k + 1)] - fwd * (vp[i] + vm[i]);
      bk[i] = bk[i] + vp[i] - vm[i];
      ck[IX(M, j, k - 1)] = ck[IX(M, j, k - 1)] + bck * (vp[i] + vm[i]);
    }
  /* diffus, src/samodel2d.f:545-688 */
  double rey = c->rey;
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js - 1; j <= je; ++j) {
      size_t i = IX(M, j, k), i1 = IX(M, j + 1, k);
      dmxh[i] = 0.5 * (xx[i] + xx[i1]);
      dmyh[i] = 0.5 * (xy[i] + xy[i1]);
      double vnulh = 0.5 * (vmul[i] / q1[i] / qj[i] + vmul[i1] / q1[i1] / qj[i1]);
      double vnuh = 0.5 * (q5[i] + q5[i1]);
      chp[i] = (1. + SA_cb2) / SA_sigma / rey * (vnulh + vnuh);
      dnuhp[i] = q5[i1] - q5[i];
    }
  for (int j = js; j <= je; ++j)
    for (int k = ks - 1; k <= ke + 1; ++k) {
      size_t i = IX(M, j, k), im1 = IX(M, j - 1, k);
      double dxp = dmxh[i] * xx[i] + dmyh[i] * xy[i];
      double dxm = dmxh[im1] * xx[i] + dmyh[im1] * xy[i];
      double c2 = SA_cb2 / SA_sigma / rey;
      c2 = c2 * (vmul[i] / q1[i] / qj[i] + q5[i]);
      double dcp = dxp * (chp[i] - c2);
      double dcm = dxm * (chp[im1] - c2);
      double ax = 0.0, cx = 0.0;
      if (k != ke + 1 && k != ks - 1) { ax = MAX(dcm, 0.0); cx = MAX(dcp, 0.0); }
      sjmat[i] = sjmat[i] - ax * dnuhp[im1] + cx * dnuhp[i];
      aj[i] = aj[i] - ax;
      cj[i] = cj[i] - cx;
      bj[i] = bj[i] + ax + cx;
    }
  for (int k = 1; k <= kd; ++k) { aj[IX(M, je + 1, k)] = 0.0; cj[IX(M, js - 1, k)] = 0.0; }
  for (int k = ks - 1; k <= ke; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k), i1 = IX(M, j, k + 1);
      dmxh[i] = 0.5 * (yx[i] + yx[i1]);
      dmyh[i] = 0.5 * (yy[i] + yy[i1]);
      double vnulh = 0.5 * (vmul[i] / q1[i] / qj[i] + vmul[i1] / q1[i1] / qj[i1]);
      double vnuh = 0.5 * (q5[i] + q5[i1]);
      chp[i] = (1. + SA_cb2) / SA_sigma / rey * (vnulh + vnuh);
      dnuhp[i] = q5[i1] - q5[i];
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k), km1 = IX(M, j, k - 1);
      double dyp = dmxh[i] * yx[i] + dmyh[i] * yy[i];
      double dym = dmxh[km1] * yx[i] + dmyh[km1] * yy[i];
      double c2 = SA_cb2 / SA_sigma / rey;
      c2 = c2 * (vmul[i] / q1[i] / qj[i] + q5[i]);
      double dcp = dyp * (chp[i] - c2);
      double dcm = dym * (chp[km1] - c2);
      double ay = 0.0, cy = 0.0;
      if (j != js - 1 && j != je + 1) { ay = MAX(dcm, 0.0); cy = MAX(dcp, 0.0); }
      sjmat[i] = sjmat[i] - ay * dnuhp[km1] + cy * dnuhp[i];
      ak[i] = ak[i] - ay;
      ck[i] = ck[i] - cy;
      bk[i] = bk[i] + ay + cy;
    }
  for (int j = 1; j <= jd; ++j) { ak[IX(M, j, ke + 1)] = 0.0; ck[IX(M, j, ks - 1)] = 0.0; }
  /* source, src/samodel2d.f:691-824 (isour = 0, fturf = 0) */
  {
    double d2min = 1.e-12, rmax = 10.0, stilim = 1.e-12, fturf = 0.0;
    double cv1_3 = ipow3(SA_cv1), cw3_6 = ipow6(SA_cw3);
    for (int k = ks; k <= ke; ++k)
      for (int j = js; j <= je; ++j) {
        size_t i = IX(M, j, k);
        double nt = q5[i];
        double vnul = vmul[i] / q1[i] / qj[i];
        double chi = nt / vnul;
        double d = sn[i];
        chi = MAX(chi, 1.e-12);
        double chi3 = ipow3(chi);
        double fv1 = chi3 / (chi3 + cv1_3);
        double fv2 = 1. - chi / (1. + chi * fv1);
        double fv3 = 1.0;
        double ft2 = fturf * SA_ct3 * exp(-1. * SA_ct4 * chi * chi);
        double dchi = 1. / vnul;
        double dfv1 = (3. * cv1_3) * ipow2(chi) * dchi * ipow2(1. / (chi3 + cv1_3));
        double dfv2 = (fv2 - 1.) / chi / vnul + (1. - fv2) * (1 - fv2) * (dfv1 + fv1 / chi / vnul);
        double dfv3 = 0.0;
        double dft2 = (-2. * SA_ct4) * chi * dchi * ft2;
        double d2 = MAX(ipow2(d), d2min);
        double stilda = vort[i] * fv3 + vnul / (d2 * SA_akt * SA_akt * rey) * chi * fv2;
        stilda = MAX(stilda, stilim);
        double r = nt / (d2 * SA_akt * SA_akt * rey) / stilda;
        r = MIN(r, rmax);
        double g = r + SA_cw2 * (ipow6(r) - r);
        double fw = (1. + cw3_6) / (ipow6(g) + cw3_6);
        fw = g * pow(fw, 1. / 6.);
        double dstild = vort[i] * dfv3 + vnul * (dchi * fv2 + chi * dfv2) / (d2 * SA_akt * SA_akt * rey);
        double dr = vnul * (dchi - chi * dstild / stilda) / stilda / (d2 * SA_akt * SA_akt * rey);
        double dg = dr * (1. + SA_cw2 * (6. * pow(r, 5.) - 1.));
        double dfw = pow((1. + cw3_6) / (ipow6(g) + cw3_6), 1. / 6.);
        dfw = dfw * dg * (1. - ipow6(g) / (ipow6(g) + cw3_6));
        double pro = SA_cb1 * stilda * (1. - ft2);
        double des = (SA_cw1 * fw - SA_cb1 / (SA_akt * SA_akt) * ft2) / d2 / rey;
        double prod = SA_cb1 * stilda * (1. - ft2) * nt;
        double dest = (SA_cw1 * fw - SA_cb1 / (SA_akt * SA_akt) * ft2) * nt * nt / d2 / rey;
        double dpro = pro * dstild / stilda - SA_cb1 * stilda * dft2;
        double ddes = (SA_cw1 * dfw - SA_cb1 / (SA_akt * SA_akt) * dft2) / d2 / rey * vnul * chi;
        ddes = ddes + des;
        ddes = ddes * nt;
        dpro = dpro * nt;
        sjmat[i] = sjmat[i] + prod - dest;
        double tk1 = MAX(des * nt - pro, 0.0);
        double tk2 = MAX(ddes - dpro, 0.0);
        bk[i] = bk[i] + tk1 + tk2;
      }
  }
  /* scale by the turbulence pseudo time step (src/samodel2d.f:101-131) */
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double tscal = tscale[i];
      if (c->p.timeac == 1) {
        tscal = (double)MAX(ib[i], 0) * (1.0 + 0.002 * sqrt(qj[i])) / (1. + sqrt(qj[i]));
        tscal = tscal * 1.0;
        tscal = tscal / (1. + tscal / h);
      }
      sjmat[i] = sjmat[i] * tscal;
      aj[i] = aj[i] * tscal; cj[i] = cj[i] * tscal; ak[i] = ak[i] * tscal; ck[i] = ck[i] * tscal;
      bjmat[i] = 1. + (bj[i] + bk[i]) * tscal;
      s1[i] = 0.; s2[i] = 0.;
    }
  /* lsolvej, src/samodel2d.f:1306-1344 */
  for (int k = ks; k <= ke; ++k) {
    size_t i = IX(M, js, k);
    double bb = 1. / bjmat[i];
    sjmat[i] = sjmat[i] * bb;
    cj[i] = cj[i] * bb;
  }
  for (int j = js + 1; j <= je; ++j)
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k), i1 = IX(M, j - 1, k);
      double bb = 1. / (bjmat[i] - aj[i] * cj[i1]);
      sjmat[i] = (sjmat[i] - aj[i] * sjmat[i1]) * bb;
      cj[i] = cj[i] * bb;
    }
  for (int j = je - 1; j >= js; --j)
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k);
      sjmat[i] = sjmat[i] - cj[i] * sjmat[IX(M, j + 1, k)];
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      s1[i] = s1[i] + sjmat[i];
      sjmat[i] = (s1[i] - s2[i]) * bjmat[i];
    }
  /* lsolvek, src/samodel2d.f:1347-1385 */
  for (int j = js; j <= je; ++j) {
    size_t i = IX(M, j, ks);
    double bb = 1. / bjmat[i];
    sjmat[i] = sjmat[i] * bb;
    ck[i] = ck[i] * bb;
  }
  for (int k = ks + 1; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k), i1 = IX(M, j, k - 1);
      double bb = 1. / (bjmat[i] - ak[i] * ck[i1]);
      sjmat[i] = (sjmat[i] - ak[i] * sjmat[i1]) * bb;
      ck[i] = ck[i] * bb;
    }
  for (int k = ke - 1; k >= ks; --k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      sjmat[i] = sjmat[i] - ck[i] * sjmat[IX(M, j, k + 1)];
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      s2[i] = s2[i] + sjmat[i];
      sjmat[i] = s2[i];
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      sjmat[i] = relfac * sjmat[i];
    }
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      q5[i] = MAX((q5[i] + sjmat[i] * (double)MAX(ib[i], 0)), vnulim);
    }
  turbc_mesh(c, M, nsp, u, v);
  /* c_turm, src/samodel2d.f:372-399 */
  for (size_t i = 0; i < n2; ++i) {
    double vnul = vmul[i] / (q1[i] * qj[i]);
    double chi = q5[i] / vnul;
    double fv1 = ipow3(chi) / (ipow3(chi) + ipow3(SA_cv1));
    turmu[i] = q1[i] * qj[i] * q5[i] * fv1;
  }
  free(sjmat); free(bjmat); free(aj); free(bj); free(cj); free(ak); free(bk); free(ck);
  free(vmul); free(vort); free(sn); free(u); free(v); free(s1); free(s2);
  free(up); free(um); free(vp); free(vm); free(chp); free(dnuhp); free(dmxh); free(dmyh);
}

/* ------------------------------------------------------- k-omega SST (a13) */
/* constants, src/komega.h:3-8 */
static const double KW_sigmak1 = 0.85, KW_sigmaw1 = 0.5, KW_beta1 = 0.075, KW_sigmak2 = 1.0,
                    KW_sigmaw2 = 0.856, KW_beta2 = 0.0828, KW_betastar = 0.09;
static const double KW_gamma1 = 5. / 9, KW_gamma2 = 0.44;

/* komegabc, src/komegasst.f:906-1434 (one halo layer).  rho, u, v, vmul
   are the arrays komegasst filled before the call. */
static int komegabc_mesh(or_case* c, or_mesh* M, int nsp, const double* rho, const double* u, const double* v,
                         const double* vmul) {
  double *q5 = QA(M, nsp, 5), *q6 = QA(M, nsp, 6);
  double *ug = GA(M, ug, nsp), *vg = GA(M, vg, nsp);
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  double *x = GA(M, x, nsp), *y = GA(M, y, nsp);
  const double *xv = VA(M, xv, nsp), *yv = VA(M, yv, nsp);
  double rei = 1. / c->rey;
  for (int ib = 0; ib < M->d.nbc; ++ib) {
    const ot2d_bcseg* b = &M->d.bc[ib];
    int js, je, ks, ke;
    bc_range(M, b, &js, &je, &ks, &ke);
    int idir = b->ibdir, iadd = idir >= 0 ? 1 : -1, iadir = abs(idir);
    int t = b->ibtyp;
    double foso = 1.0;
    if (t == 3 || t == 4) {          /* komegabc_extpt, 1079-1129 */
      if (iadir == 1) {
        for (int jc = 1; jc <= je - js + 1; ++jc) {
          int j = (idir == 1) ? je - jc + 1 : js + jc - 1;
          int j1 = j + iadd, j2 = j1 + iadd;
          for (int k = ks; k <= ke; ++k) {
            q5[IX(M, j, k)] = (1. + foso) * q5[IX(M, j1, k)] - foso * q5[IX(M, j2, k)];
            q6[IX(M, j, k)] = (1. + foso) * q6[IX(M, j1, k)] - foso * q6[IX(M, j2, k)];
          }
        }
      } else if (iadir == 2) {
        for (int kc = 1; kc <= ke - ks + 1; ++kc) {
          int k = (idir == 2) ? ke - kc + 1 : ks + kc - 1;
          int k1 = k + iadd, k2 = k1 + iadd;
          for (int j = js; j <= je; ++j) {
            q5[IX(M, j, k)] = (1. + foso) * q5[IX(M, j, k1)] - foso * q5[IX(M, j, k2)];
            q6[IX(M, j, k)] = (1. + foso) * q6[IX(M, j, k1)] - foso * q6[IX(M, j, k2)];
          }
        }
      }
    } else if (t == 5) {             /* komegabc_wall, 977-1076 (one halo layer) */
      if (iadir != 2) return OT2D_ERR_UNSUPPORTED;
      double tt = ((double)c->istep0 - 1.) / 30.;
      if (tt > 1. || c->p.iread > 0) tt = 1.;
      double scal = (10. - 15. * tt + 6. * tt * tt) * ((tt * tt) * tt);
      int k = (idir == 2) ? ke : ks, k1 = k + iadd, kv = k1 - 1;
      for (int j = js; j <= je; ++j) {
        int jv = j - 1;
        size_t i = IX(M, j, k), i1 = IX(M, j, k1);
        double dely2 = sq(x[i1] - 0.5 * (xv[VX(M, jv, kv)] + xv[VX(M, jv + 1, kv)])) +
                       sq(y[i1] - 0.5 * (yv[VX(M, jv, kv)] + yv[VX(M, jv + 1, kv)]));
        double vmulwall = 0.5 * (vmul[i1] + vmul[i1]);
        double omegawall = 60. * vmulwall * rei * rei / (KW_beta1 * dely2);
        double qwall5 = -q5[i1];
        double qwall6 = 2 * omegawall - 6. * vmul[i1] * rei * rei / (KW_beta1 * dely2);
        q5[i] = scal * qwall5 + (1. - scal) * q5[i];
        q6[i] = scal * qwall6 + (1. - scal) * q6[i];
      }
    } else if (t == 10 || t == 11) { /* komegabc_sym, 1132-1180 */
      if (iadir == 1) {
        for (int jc = 1; jc <= je - js + 1; ++jc) {
          int j, j1;
          if (idir == 1) { j = je - jc + 1; j1 = je + jc; } else { j = js + jc - 1; j1 = js - jc; }
          for (int k = ks; k <= ke; ++k) { q5[IX(M, j, k)] = q5[IX(M, j1, k)]; q6[IX(M, j, k)] = q6[IX(M, j1, k)]; }
        }
      } else if (iadir == 2) {
        for (int kc = 1; kc <= ke - ks + 1; ++kc) {
          int k, k1;
          if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
          for (int j = js; j <= je; ++j) { q5[IX(M, j, k)] = q5[IX(M, j, k1)]; q6[IX(M, j, k)] = q6[IX(M, j, k1)]; }
        }
      }
    } else if (t == 22) {            /* komegabc_periodic, 1183-1251 */
      if (iadir == 1) {
        for (int j = js; j <= je; ++j) {
          int jc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, j);
          for (int k = ks; k <= ke; ++k) { q5[IX(M, j, k)] = q5[IX(M, jc, k)]; q6[IX(M, j, k)] = q6[IX(M, jc, k)]; }
        }
      } else {
        for (int k = ks; k <= ke; ++k) {
          int kc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, k);
          for (int j = js; j <= je; ++j) { q5[IX(M, j, k)] = q5[IX(M, j, kc)]; q6[IX(M, j, k)] = q6[IX(M, j, kc)]; }
        }
      }
    } else if (t == 51 && iadir == 2) { /* komegabc_wake, 1254-1295 */
      for (int kc = 1; kc <= ke - ks + 1; ++kc) {
        int k, k1;
        if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
        for (int j = js; j <= je; ++j) {
          int jj = M->jd - j + 1;
          q5[IX(M, j, k)] = q5[IX(M, jj, k1)];
          q6[IX(M, j, k)] = q6[IX(M, jj, k1)];
          q5[IX(M, jj, k)] = q5[IX(M, j, k1)];
          q6[IX(M, jj, k)] = q6[IX(M, j, k1)];
        }
      }
    } else if (t == 46) {            /* komegabc_inf, 1298-1332 */
      for (int k = ks; k <= ke; ++k)
        for (int j = js; j <= je; ++j) {
          size_t i = IX(M, j, k);
          q5[i] = c->tkeinf * rho[i];
          q6[i] = c->tomegainf * rho[i];
        }
    } else if (t == 47) {            /* komegabc_out, 1335-1432 */
      if (iadir == 1) {
        int j = (idir == 1) ? je : js, j1 = j + iadd;
        for (int k = ks; k <= ke; ++k) {
          size_t i = IX(M, j, k), i1 = IX(M, j1, k);
          q5[i] = c->tkeinf * rho[i];
          q6[i] = c->tomegainf * rho[i];
          double uu = (u[i] - ug[i] + u[i1] - ug[i1]) * xx[i];
          uu = uu + (v[i] - vg[i] + v[i1] - vg[i1]) * xy[i];
          uu = uu * iadd;
          if (uu < 0.) { q5[i] = q5[i1]; q6[i] = q6[i1]; }
        }
      } else if (iadir == 2) {
        int k = (idir == 2) ? ke : ks, k1 = k + iadd;
        for (int j = js; j <= je; ++j) {
          size_t i = IX(M, j, k), i1 = IX(M, j, k1);
          q5[i] = c->tkeinf * rho[i];
          q6[i] = c->tomegainf * rho[i];
          double uu = (u[i] - ug[i] + u[i1] - ug[i1]) * yx[i];
          uu = uu + (v[i] - vg[i] + v[i1] - vg[i1]) * yy[i];
          uu = uu * iadd;
          if (uu < 0.) { q5[i] = q5[i1]; q6[i] = q6[i1]; }
        }
      }
    }
  }
  return OT2D_OK;
}

/* tridiagonal solves of one equation along j (lsolvej2, 1435-1475) or
   along k (lsolvek2, 1478-1518): stride st between line cells, n cells */
static void kw_thomas(double* a, const double* b, double* cc, double* s, size_t i0, size_t st, int n) {
  double bb = 1. / b[i0];
  s[i0] = s[i0] * bb;
  cc[i0] = cc[i0] * bb;
  for (int t = 1; t < n; ++t) {
    size_t i = i0 + t * st, im = i - st;
    bb = 1. / (b[i] - a[i] * cc[im]);
    s[i] = (s[i] - a[i] * s[im]) * bb;
    cc[i] = cc[i] * bb;
  }
  for (int t = n - 2; t >= 0; --t) {
    size_t i = i0 + t * st;
    s[i] = s[i] - cc[i] * s[i + st];
  }
}

/* komegasst, src/komegasst.f:2-187 (itrans = 0).  The reference reads F1,
   F2 and strain on the halo ring without writing them there (blend and
   kwvortic fill the interior only; findturm and komegadiffus read every
   cell): allocations here are zero-filled, and the reference run the same
   way (zero-filled malloc) reproduces these values bit for bit. */
static int komegasst_mesh(or_case* c, or_mesh* M, int nsp) {
  int jd = M->jd, kd = M->kd, js = M->jbeg, je = M->jend, ks = M->kbeg, ke = M->kend;
  size_t n2 = M->n2;
  double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
  double *q5 = QA(M, nsp, 5), *q6 = QA(M, nsp, 6), *qj = QA(M, nsp, c->nq);
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  double *ug = GA(M, ug, nsp), *vg = GA(M, vg, nsp);
  double *x = GA(M, x, nsp), *y = GA(M, y, nsp);
  const double *xv = VA(M, xv, nsp), *yv = VA(M, yv, nsp);
  double* turmu = GA(M, turmu, nsp);
  double* tscale = GA(M, tscale, nsp);
  const int32_t* ib = M->iblank + (size_t)nsp * n2;
  double *s5 = SA(M, nsp, 5), *s6 = SA(M, nsp, 6);
  double *rho = calloc(n2, 8), *u = calloc(n2, 8), *v = calloc(n2, 8), *vmul = calloc(n2, 8), *sn = calloc(n2, 8);
  double *tauxx = calloc(n2, 8), *tauxy = calloc(n2, 8), *tauyy = calloc(n2, 8);
  double *ux = calloc(n2, 8), *uy = calloc(n2, 8), *vx = calloc(n2, 8), *vy = calloc(n2, 8);
  double *cdkw = calloc(n2, 8), *F1 = calloc(n2, 8), *F2 = calloc(n2, 8), *strain = calloc(n2, 8);
  double *aj[2], *bj[2], *cj[2], *ak[2], *ck[2], *sj[2];
  for (int n = 0; n < 2; ++n) {
    aj[n] = calloc(n2, 8); bj[n] = calloc(n2, 8); cj[n] = calloc(n2, 8);
    ak[n] = calloc(n2, 8); ck[n] = calloc(n2, 8); sj[n] = calloc(n2, 8);
  }
  double *chp = calloc(n2, 8), *dnuhp = calloc(n2, 8), *dmxh = calloc(n2, 8), *dmyh = calloc(n2, 8);
  double *up = calloc(n2, 8), *um = calloc(n2, 8);
  double* qk[2] = {q5, q6};
  double* sk[2] = {s5, s6};
  double rei = 1. / c->rey;
  int rc = OT2D_OK;
  /* lamvis, src/samodel2d.f:200-238 */
  {
    double c2b = 198.6 / c->p.tinf, c2bp = c2b + 1.;
    for (size_t i = 0; i < n2; ++i) {
      double ra = 1. / q1[i];
      double uvel = q2[i] * ra, vvel = q3[i] * ra;
      double ei = q4[i] * ra - 0.5 * (sq(uvel) + sq(vvel));
      double tt = c->ggm1 * ei;
      vmul[i] = c2bp * tt * sqrt(tt) / (c2b + tt);
    }
  }
  for (size_t i = 0; i < n2; ++i) {
    rho[i] = q1[i] * qj[i];
    u[i] = q2[i] / q1[i];
    v[i] = q3[i] / q1[i];
    sn[i] = 1.e10;
  }
  if ((rc = komegabc_mesh(c, M, nsp, rho, u, v, vmul))) goto done;
  /* kwdist, 218-267 */
  if (M->bodyflag) {
    for (int j = M->jtail1; j <= M->jtail2; ++j) sn[IX(M, j, 1)] = 0.0;
    for (int j = M->jtail1; j <= M->jtail2; ++j)
      for (int k = 2; k <= kd; ++k) {
        int jv = j - 1;
        size_t i = IX(M, j, k);
        sn[i] = sqrt(sq(x[i] - 0.5 * (xv[VX(M, jv, 1)] + xv[VX(M, jv + 1, 1)])) +
                     sq(y[i] - 0.5 * (yv[VX(M, jv, 1)] + yv[VX(M, jv + 1, 1)])));
      }
    int jv = M->jtail1 - 1;
    for (int j = 1; j <= M->jtail1 - 1; ++j)
      for (int k = 1; k <= kd; ++k) {
        size_t i = IX(M, j, k);
        sn[i] = sqrt(sq(x[i] - xv[VX(M, jv, 1)]) + sq(y[i] - yv[VX(M, jv, 1)]));
      }
    for (int j = M->jtail2 + 1; j <= jd; ++j)
      for (int k = 1; k <= kd; ++k) {
        size_t i = IX(M, j, k);
        sn[i] = sqrt(sq(x[i] - xv[VX(M, jv, 1)]) + sq(y[i] - yv[VX(M, jv, 1)]));
      }
  }
  /* kwvortic (270-318, strain only), turbstress (321-371), crossdiff
     (374-414) and blend (417-461) over the interior */
  for (int j = js; j <= je; ++j)
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k);
      size_t ie = IX(M, j + 1, k), iw = IX(M, j - 1, k), in = IX(M, j, k + 1), is = IX(M, j, k - 1);
      double usi = 0.5 * (u[ie] - u[iw]), vsi = 0.5 * (v[ie] - v[iw]);
      double ueta = 0.5 * (u[in] - u[is]), veta = 0.5 * (v[in] - v[is]);
      double lux = xx[i] * usi + yx[i] * ueta, luy = xy[i] * usi + yy[i] * ueta;
      double lvx = xx[i] * vsi + yx[i] * veta, lvy = xy[i] * vsi + yy[i] * veta;
      double sxx = lux, sxy = 0.5 * (luy + lvx), syy = lvy;
      strain[i] = sqrt(2. * (sxx * sxx + 2. * sxy * sxy + syy * syy));
      ux[i] = lux; uy[i] = luy; vx[i] = lvx; vy[i] = lvy;
      double sfdiv = 2. / 3 * (ux[i] + vy[i]);
      tauxx[i] = turmu[i] * rei * (2.0 * ux[i] - sfdiv) - 2. / 3 * q5[i];
      tauyy[i] = turmu[i] * rei * (2.0 * vy[i] - sfdiv) - 2. / 3 * q5[i];
      tauxy[i] = turmu[i] * rei * (uy[i] + vx[i]);
    }
  for (int j = js; j <= je; ++j)
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k);
      size_t ie = IX(M, j + 1, k), iw = IX(M, j - 1, k), in = IX(M, j, k + 1), is = IX(M, j, k - 1);
      double ksi = 0.5 * (q5[ie] / rho[ie] - q5[iw] / rho[iw]);
      double keta = 0.5 * (q5[in] / rho[in] - q5[is] / rho[is]);
      double omegasi = 0.5 * (q6[ie] / rho[ie] - q6[iw] / rho[iw]);
      double omegaeta = 0.5 * (q6[in] / rho[in] - q6[is] / rho[is]);
      double kx = ksi * xx[i] + keta * yx[i], ky = ksi * xy[i] + keta * yy[i];
      double omegax = omegasi * xx[i] + omegaeta * yx[i], omegay = omegasi * xy[i] + omegaeta * yy[i];
      cdkw[i] = kx * omegax + ky * omegay;
      cdkw[i] = cdkw[i] * 2 * KW_sigmaw2 * rho[i] * rho[i] / q6[i];
    }
  for (int j = js; j <= je; ++j)
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k);
      double a1 = rei * sqrt(q5[i] * rho[i]) / (0.09 * q6[i] * sn[i]);
      double a2 = 500 * rei * rei * vmul[i] / (q6[i] * sq(sn[i]));
      double cd = MAX(cdkw[i], 1e-10);
      double a3 = 4 * q5[i] * KW_sigmaw2 / (cd * sq(sn[i]));
      double arg1 = MIN(MAX(a1, a2), a3);
      double arg2 = MAX(2 * a1, a2);
      F1[i] = tanh(((arg1 * arg1) * arg1) * arg1);
      F2[i] = tanh(arg2 * arg2);
    }
  /* findturm, 190-215 */
  for (size_t i = 0; i < n2; ++i) {
    double a1 = 0.31;
    turmu[i] = a1 * rho[i] * q5[i] / MAX(a1 * q6[i], F2[i] * rho[i] * strain[i] * rei);
    turmu[i] = MAX(turmu[i], 1e-20);
  }
  /* komegarhslhs, 464-590 */
  for (size_t i = 0; i < n2; ++i)
    for (int n = 0; n < 2; ++n) { aj[n][i] = bj[n][i] = cj[n][i] = ak[n][i] = ck[n][i] = 0.0; sj[n][i] = sk[n][i]; }
  for (int k = ks; k <= ke; ++k) {
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      double uu = xx[i] * (u[i] - ug[i]) + xy[i] * (v[i] - vg[i]);
      up[i] = 0.5 * (uu + fabs(uu));
      um[i] = 0.5 * (uu - fabs(uu));
    }
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k), iw = IX(M, j - 1, k), ie = IX(M, j + 1, k);
      for (int n = 0; n < 2; ++n)
        sj[n][i] = sj[n][i] - up[i] * (qk[n][i] - qk[n][iw]) - um[i] * (qk[n][ie] - qk[n][i]);
      for (int n = 0; n < 2; ++n) {
        aj[n][i] = aj[n][i] - up[i];
        cj[n][i] = cj[n][i] + um[i];
        bj[n][i] = bj[n][i] + up[i] - um[i];
      }
    }
  }
  for (int j = js; j <= je; ++j) {
    for (int k = ks - 1; k <= ke + 1; ++k) {
      size_t i = IX(M, j, k);
      double vv = yx[i] * (u[i] - ug[i]) + yy[i] * (v[i] - vg[i]);
      up[i] = 0.5 * (vv + fabs(vv));
      um[i] = 0.5 * (vv - fabs(vv));
    }
    for (int k = ks; k <= ke; ++k) {
      size_t i = IX(M, j, k), is = IX(M, j, k - 1), in = IX(M, j, k + 1);
      for (int n = 0; n < 2; ++n)
        sj[n][i] = sj[n][i] - up[i] * (qk[n][i] - qk[n][is]) - um[i] * (qk[n][in] - qk[n][i]);
      for (int n = 0; n < 2; ++n) {
        ak[n][i] = ak[n][i] - up[i];
        ck[n][i] = ck[n][i] + um[i];
        bj[n][i] = bj[n][i] + up[i] - um[i];
      }
    }
  }
  /* komegasource, 847-903 (itrans = 0) */
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double prod = tauxx[i] * ux[i] + tauxy[i] * (uy[i] + vx[i]) + tauyy[i] * vy[i];
      double kdes = KW_betastar * c->rey * q6[i] * q5[i] / rho[i];
      double ksust = KW_betastar * c->rey * c->tkeinf * c->tomegainf * rho[i];
      double prodlim = 10. * kdes;
      prod = MIN(prod, prodlim);
      double kprod = prod;
      double beta = F1[i] * KW_beta1 + (1 - F1[i]) * KW_beta2;
      double omegades = beta * c->rey * q6[i] * q6[i] / rho[i];
      double omegasust = beta * c->rey * (c->tomegainf * c->tomegainf) * rho[i];
      double omegacrossdiff = (1 - F1[i]) * cdkw[i] * rei;
      sj[0][i] = sj[0][i] + kprod - kdes + ksust;
      double gammap = F1[i] * KW_gamma1 + (1 - F1[i]) * KW_gamma2;
      sj[1][i] = sj[1][i] + gammap * rho[i] / turmu[i] * prod - omegades + omegacrossdiff + omegasust;
      bj[0][i] = bj[0][i] + kdes / q5[i];
      bj[1][i] = bj[1][i] + (2 * omegades + fabs(omegacrossdiff)) / q6[i];
    }
  /* komegadiffus, 593-844 */
  for (int j = js - 1; j <= je; ++j)
    for (int k = ks - 1; k <= ke + 1; ++k) {
      size_t i = IX(M, j, k), i1 = IX(M, j + 1, k);
      dmxh[i] = 0.5 * (xx[i] + xx[i1]);
      dmyh[i] = 0.5 * (xy[i] + xy[i1]);
    }
  for (int n = 0; n < 2; ++n) {
    double s1 = n == 0 ? KW_sigmak1 : KW_sigmaw1, s2 = n == 0 ? KW_sigmak2 : KW_sigmaw2;
    for (int j = js - 1; j <= je; ++j)
      for (int k = ks - 1; k <= ke + 1; ++k) {
        size_t i = IX(M, j, k), i1 = IX(M, j + 1, k);
        double vnulh = 0.5 * (vmul[i] + vmul[i1]);
        double vnuh = 0.5 * (turmu[i] + turmu[i1]);
        double sig = F1[i] * s1 + (1. - F1[i]) * s2;
        chp[i] = rei * (vnulh + sig * vnuh);
        dnuhp[i] = qk[n][i1] / rho[i1] - qk[n][i] / rho[i];
      }
    for (int j = js; j <= je; ++j)
      for (int k = ks - 1; k <= ke + 1; ++k) {
        size_t i = IX(M, j, k), iw = IX(M, j - 1, k), ie = IX(M, j + 1, k);
        double dxp = dmxh[i] * xx[i] + dmyh[i] * xy[i];
        double dxm = dmxh[iw] * xx[i] + dmyh[iw] * xy[i];
        double dcp = dxp * chp[i], dcm = dxm * chp[iw];
        double ax = 0.0, cx = 0.0;
        if (k != ke + 1 && k != ks - 1) { ax = MAX(dcm, 0.0); cx = MAX(dcp, 0.0); }
        sj[n][i] = sj[n][i] - ax * dnuhp[iw] + cx * dnuhp[i];
        aj[n][i] = aj[n][i] - ax / rho[iw];
        cj[n][i] = cj[n][i] - cx / rho[ie];
        bj[n][i] = bj[n][i] + (ax + cx) / rho[i];
      }
    for (int k = 1; k <= kd; ++k) { aj[n][IX(M, je + 1, k)] = 0.0; cj[n][IX(M, js - 1, k)] = 0.0; }
  }
  for (int k = ks - 1; k <= ke; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k), i1 = IX(M, j, k + 1);
      dmxh[i] = 0.5 * (yx[i] + yx[i1]);
      dmyh[i] = 0.5 * (yy[i] + yy[i1]);
    }
  for (int n = 0; n < 2; ++n) {
    double s1 = n == 0 ? KW_sigmak1 : KW_sigmaw1, s2 = n == 0 ? KW_sigmak2 : KW_sigmaw2;
    for (int k = ks - 1; k <= ke; ++k)
      for (int j = js - 1; j <= je + 1; ++j) {
        size_t i = IX(M, j, k), i1 = IX(M, j, k + 1);
        double vnulh = 0.5 * (vmul[i] + vmul[i1]);
        double vnuh = 0.5 * (turmu[i] + turmu[i1]);
        double sig = F1[i] * s1 + (1. - F1[i]) * s2;
        chp[i] = rei * (vnulh + sig * vnuh);
        dnuhp[i] = qk[n][i1] / rho[i1] - qk[n][i] / rho[i];
      }
    for (int k = ks; k <= ke; ++k)
      for (int j = js - 1; j <= je + 1; ++j) {
        size_t i = IX(M, j, k), is = IX(M, j, k - 1), in = IX(M, j, k + 1);
        double dyp = dmxh[i] * yx[i] + dmyh[i] * yy[i];
        double dym = dmxh[is] * yx[i] + dmyh[is] * yy[i];
        double dcp = dyp * chp[i], dcm = dym * chp[is];
        double ay = 0.0, cy = 0.0;
        if (j != js - 1 && j != je + 1) { ay = MAX(dcm, 0.0); cy = MAX(dcp, 0.0); }
        sj[n][i] = sj[n][i] - ay * dnuhp[is] + cy * dnuhp[i];
        ak[n][i] = ak[n][i] - ay / rho[is];
        ck[n][i] = ck[n][i] - cy / rho[in];
        bj[n][i] = bj[n][i] + (ay + cy) / rho[i];
      }
    for (int j = 1; j <= kd && j <= jd; ++j) { ak[n][IX(M, j, ke + 1)] = 0.0; ck[n][IX(M, j, ks - 1)] = 0.0; }
  }
  /* pseudo-time scaling, 104-131 (oat = 1) */
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double tscal = tscale[i];
      if (c->p.timeac == 1.) {
        double dtpseudo_turb = 10.0, oat = 1.0;
        tscal = (double)MAX(ib[i], 0) * (1.0 + 0.002 * sqrt(qj[i])) / (1. + sqrt(qj[i]));
        tscal = tscal * dtpseudo_turb;
        tscal = tscal / (1. + tscal / c->h / oat);
      }
      for (int n = 0; n < 2; ++n) {
        aj[n][i] = aj[n][i] * tscal;
        cj[n][i] = cj[n][i] * tscal;
        ak[n][i] = ak[n][i] * tscal;
        ck[n][i] = ck[n][i] * tscal;
        bj[n][i] = 1. + bj[n][i] * tscal;
        sj[n][i] = sj[n][i] * tscal;
      }
    }
  for (int n = 0; n < 2; ++n)
    for (int k = ks; k <= ke; ++k) kw_thomas(aj[n], bj[n], cj[n], sj[n], IX(M, js, k), 1, je - js + 1);
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j)
      for (int n = 0; n < 2; ++n) sj[n][IX(M, j, k)] = sj[n][IX(M, j, k)] * bj[n][IX(M, j, k)];
  for (int n = 0; n < 2; ++n)
    for (int j = js; j <= je; ++j) kw_thomas(ak[n], bj[n], ck[n], sj[n], IX(M, j, ks), (size_t)jd, ke - ks + 1);
  /* update with the clip (relfac = 1), 146-172 */
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      size_t i = IX(M, j, k);
      double ibp = (double)MAX(ib[i], 0);
      s5[i] = 1. * sj[0][i];
      s6[i] = 1. * sj[1][i];
      q5[i] = q5[i] + s5[i] * ibp;
      q6[i] = q6[i] + s6[i] * ibp;
      q5[i] = MAX(q5[i], 1e-20 * rho[i]);
      q6[i] = MAX(q6[i], 1e-20 * rho[i]);
    }
  if ((rc = komegabc_mesh(c, M, nsp, rho, u, v, vmul))) goto done;
  for (size_t i = 0; i < n2; ++i) {
    double a1 = 0.31;
    turmu[i] = a1 * rho[i] * q5[i] / MAX(a1 * q6[i], F2[i] * rho[i] * strain[i] * rei);
    turmu[i] = MAX(turmu[i], 1e-20);
  }
done:
  free(rho); free(u); free(v); free(vmul); free(sn); free(tauxx); free(tauxy); free(tauyy);
  free(ux); free(uy); free(vx); free(vy); free(cdkw); free(F1); free(F2); free(strain);
  for (int n = 0; n < 2; ++n) { free(aj[n]); free(bj[n]); free(cj[n]); free(ak[n]); free(ck[n]); free(sj[n]); }
  free(chp); free(dnuhp); free(dmxh); free(dmyh); free(up); free(um);
  return rc;
}

/* ------------------------------------------------ overset interpolation (a19) */
/* connectivity of mesh im, instance nsp, in the layout connect2d returns
   (src/Turns_overset.f:223-226): iblank (jd*kd, after update_halo_iblanks),
   imesh(nfringe,2), ibc(nfringe), idonor(ndonor,2), frac(ndonor,2) */
int32_t or_set_mesh_connectivity(or_case* c, int32_t nsp, int32_t im, const int32_t* iblank, int32_t nfringe,
                                 const int32_t* imesh, const int32_t* ibc, int32_t ndonor, const int32_t* idonor,
                                 const double* frac) {
  if (!c || im < 0 || im >= c->nmesh || nsp < 0 || nsp >= c->nspec || nfringe < 0 || ndonor < 0)
    return OT2D_ERR_ARG;
  or_mesh* M = &c->m[im];
  if (!M->nfringe) {
    M->nfringe = calloc(c->nspec, sizeof(int32_t));
    M->ndonor = calloc(c->nspec, sizeof(int32_t));
    M->imesh = calloc(c->nspec, sizeof(int32_t*));
    M->ibc = calloc(c->nspec, sizeof(int32_t*));
    M->idonor = calloc(c->nspec, sizeof(int32_t*));
    M->frac = calloc(c->nspec, sizeof(double*));
  }
  free(M->imesh[nsp]); free(M->ibc[nsp]); free(M->idonor[nsp]); free(M->frac[nsp]);
  M->nfringe[nsp] = nfringe;
  M->ndonor[nsp] = ndonor;
  M->imesh[nsp] = malloc(sizeof(int32_t) * (2 * (size_t)nfringe + 1));
  M->ibc[nsp] = malloc(sizeof(int32_t) * ((size_t)nfringe + 1));
  M->idonor[nsp] = malloc(sizeof(int32_t) * (2 * (size_t)ndonor + 1));
  M->frac[nsp] = malloc(sizeof(double) * (2 * (size_t)ndonor + 1));
  memcpy(M->imesh[nsp], imesh, sizeof(int32_t) * 2 * (size_t)nfringe);
  memcpy(M->ibc[nsp], ibc, sizeof(int32_t) * (size_t)nfringe);
  memcpy(M->idonor[nsp], idonor, sizeof(int32_t) * 2 * (size_t)ndonor);
  memcpy(M->frac[nsp], frac, sizeof(double) * 2 * (size_t)ndonor);
  memcpy(M->iblank + (size_t)nsp * M->n2, iblank, sizeof(int32_t) * M->n2);
  c->overset = 1;
  return OT2D_OK;
}

/* do_interpolations, src/connectivity.f:2-189 */
static void do_interpolations(or_case* c) {
  double onefourth = 1. / 4;
  for (int nsp = 0; nsp < c->nspec; ++nsp) {
    int bcdim = 0;
    for (int im = 0; im < c->nmesh; ++im) bcdim += c->m[im].nfringe ? c->m[im].ndonor[nsp] : 0;
    double* qbc = calloc((size_t)(bcdim > 0 ? bcdim : 1) * c->nv, sizeof(double));
    int iisptr = 1;
    for (int im = 0; im < c->nmesh; ++im) {
      or_mesh* M = &c->m[im];
      if (!M->nfringe) continue;
      const int nd = M->ndonor[nsp];
      const double* qj = QA(M, nsp, c->nq);
      for (int id = 1; id <= nd; ++id) {
        int ii = M->idonor[nsp][id - 1], jj = M->idonor[nsp][nd + id - 1];
        int iim = ii - 1, jjm = jj - 1, iip = ii + 1, jjp = jj + 1;
        double f1 = M->frac[nsp][id - 1], f2 = M->frac[nsp][nd + id - 1];
        double djm1 = 1. + f1, dkm1 = 1. + f2, dj0 = f1, dk0 = f2, djp1 = 1. - f1, dkp1 = 1. - f2;
        double w1 = (dj0 * djp1 * dk0 * dkp1) * onefourth;
        double w2 = -2 * (djm1 * djp1 * dk0 * dkp1) * onefourth;
        double w3 = -(djm1 * dj0 * dk0 * dkp1) * onefourth;
        double w4 = -2 * (dj0 * djp1 * dkm1 * dkp1) * onefourth;
        double w5 = 4 * (djm1 * djp1 * dkm1 * dkp1) * onefourth;
        double w6 = 2 * (djm1 * dj0 * dkm1 * dkp1) * onefourth;
        double w7 = -(dj0 * djp1 * dkm1 * dk0) * onefourth;
        double w8 = 2 * (djm1 * djp1 * dkm1 * dk0) * onefourth;
        double w9 = (djm1 * dj0 * dkm1 * dk0) * onefourth;
        size_t a = IX(M, iim, jjm), b = IX(M, ii, jjm), cc = IX(M, iip, jjm);
        size_t d = IX(M, iim, jj), e = IX(M, ii, jj), f = IX(M, iip, jj);
        size_t g = IX(M, iim, jjp), h = IX(M, ii, jjp), k = IX(M, iip, jjp);
        double* row = qbc + (size_t)(iisptr - 1 + id - 1) * c->nv;
        for (int n = 1; n <= c->nmv; ++n) {
          const double* q = QA(M, nsp, n);
          row[n - 1] = w1 * q[a] * qj[a] + w2 * q[b] * qj[b] + w3 * q[cc] * qj[cc] + w4 * q[d] * qj[d] +
                       w5 * q[e] * qj[e] + w6 * q[f] * qj[f] + w7 * q[g] * qj[g] + w8 * q[h] * qj[h] +
                       w9 * q[k] * qj[k];
        }
        for (int n = c->nmv + 1; n <= c->nv; ++n) {
          const double* q = QA(M, nsp, n);
          row[n - 1] = w1 * q[a] + w2 * q[b] + w3 * q[cc] + w4 * q[d] + w5 * q[e] + w6 * q[f] + w7 * q[g] +
                       w8 * q[h] + w9 * q[k];
        }
      }
      iisptr += nd;
    }
    for (int im = 0; im < c->nmesh; ++im) {
      or_mesh* M = &c->m[im];
      if (!M->nfringe) continue;
      const int nf = M->nfringe[nsp];
      double* qj = QA(M, nsp, c->nq);
      for (int id = 1; id <= nf; ++id) {
        int j = M->ibc[nsp][id - 1];
        int ii = M->imesh[nsp][id - 1], jj = M->imesh[nsp][nf + id - 1];
        size_t i = IX(M, ii, jj);
        const double* row = qbc + (size_t)(j - 1) * c->nv;
        for (int n = 1; n <= c->nmv; ++n) QA(M, nsp, n)[i] = row[n - 1] / qj[i];
        for (int n = c->nmv + 1; n <= c->nv; ++n) QA(M, nsp, n)[i] = row[n - 1];
      }
    }
    free(qbc);
  }
}

int32_t or_interpolate(or_case* c) {
  if (!c) return OT2D_ERR_ARG;
  if (c->overset) do_interpolations(c);
  return OT2D_OK;
}

/* compute_residual, src/step.f:221-300 */
int32_t or_compute_residual(or_case* c) {
  for (int im = 0; im < c->nmesh; ++im) {
    or_mesh* M = &c->m[im];
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      bc_mesh(c, M, nsp);
      int rc = rhsup_mesh(c, M, nsp);
      if (rc) return rc;
      if (!c->p.invisc) visrhs_mesh(c, M, nsp);
      if (c->timespectral) timespectralrhs(c, M, nsp);
    }
  }
  return OT2D_OK;
}

/* solve_turbulence_equations, src/step.f:303-329 (SA and SST, no transition) */
int32_t or_turbulence(or_case* c) {
  if (c->lamin) return OT2D_OK;
  if ((c->iturb != 1 && c->iturb != 2) || c->p.itrans) return OT2D_ERR_UNSUPPORTED;
  for (int im = 0; im < c->nmesh; ++im)
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      if (c->iturb == 1) vmu_sa_mesh(c, &c->m[im], nsp);
      else {
        int rc = komegasst_mesh(c, &c->m[im], nsp);
        if (rc) return rc;
      }
    }
  return OT2D_OK;
}

/* ---------------------------------------------------- residual BC (a7) */
/* rbc + zero_residuals + rbcwake/rbcextp/rbcsym/rbcperiodic,
 * src/bc.f:88-177, 1636-1666, 1576-1633, 2029-2094, 2160-2222, 270-357 */
static void rbc_mesh(or_case* c, or_mesh* M, int nsp) {
  double* qj = QA(M, nsp, c->nq);
  double* ss[4] = {SA(M, nsp, 1), SA(M, nsp, 2), SA(M, nsp, 3), SA(M, nsp, 4)};
  for (int ib = 0; ib < M->d.nbc; ++ib) {
    const ot2d_bcseg* b = &M->d.bc[ib];
    int js, je, ks, ke;
    bc_range(M, b, &js, &je, &ks, &ke);
    int idir = b->ibdir, iadd = idir >= 0 ? 1 : -1, iadir = abs(idir);
    int t = b->ibtyp;
    if (t == 3 || t == 4 || t == 5 || t == 46 || t == 47) {
      for (int j = js; j <= je; ++j)
        for (int k = ks; k <= ke; ++k)
          for (int n = 0; n < 4; ++n) ss[n][IX(M, j, k)] = 0.0;
    } else if (t == 10) {
      double foso = 0.0;
      if (iadir == 1) {
        for (int jc = 1; jc <= je - js + 1; ++jc) {
          int j = (idir == 1) ? je - jc + 1 : js + jc - 1;
          int j1 = j + iadd, j2 = j1 + iadd;
          for (int k = ks; k <= ke; ++k) {
            size_t i = IX(M, j, k), i1 = IX(M, j1, k), i2 = IX(M, j2, k);
            double s1 = qj[i1] / qj[i], s2 = qj[i2] / qj[i];
            for (int n = 0; n < 4; ++n) ss[n][i] = (1. + foso) * ss[n][i1] * s1 - foso * ss[n][i2] * s2;
          }
        }
      } else if (iadir == 2) {
        for (int kc = 1; kc <= ke - ks + 1; ++kc) {
          int k = (idir == 2) ? ke - kc + 1 : ks + kc - 1;
          int k1 = k + iadd, k2 = k1 + iadd;
          for (int j = js; j <= je; ++j) {
            size_t i = IX(M, j, k), i1 = IX(M, j, k1), i2 = IX(M, j, k2);
            double s1 = qj[i1] / qj[i], s2 = qj[i2] / qj[i];
            for (int n = 0; n < 4; ++n) ss[n][i] = (1. + foso) * ss[n][i1] * s1 - foso * ss[n][i2] * s2;
          }
        }
      }
    } else if (t == 11) {
      if (iadir == 1) {
        for (int jc = 1; jc <= je - js + 1; ++jc) {
          int j, j1;
          if (idir == 1) { j = je - jc + 1; j1 = je + jc; } else { j = js + jc - 1; j1 = js - jc; }
          for (int k = ks; k <= ke; ++k) {
            size_t i = IX(M, j, k), i1 = IX(M, j1, k);
            double sc = qj[i1] / qj[i];
            ss[0][i] = ss[0][i1] * sc; ss[1][i] = -ss[1][i1] * sc;
            ss[2][i] = ss[2][i1] * sc; ss[3][i] = ss[3][i1] * sc;
          }
        }
      } else if (iadir == 2) {
        for (int kc = 1; kc <= ke - ks + 1; ++kc) {
          int k, k1;
          if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
          for (int j = js; j <= je; ++j) {
            size_t i = IX(M, j, k), i1 = IX(M, j, k1);
            double sc = qj[i1] / qj[i];
            ss[0][i] = ss[0][i1] * sc; ss[1][i] = ss[1][i1] * sc;
            ss[2][i] = -ss[2][i1] * sc; ss[3][i] = ss[3][i1] * sc;
          }
        }
      }
    } else if (t == 22) {
      if (iadir == 1) {
        for (int j = js; j <= je; ++j) {
          int jc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, j);
          for (int k = ks; k <= ke; ++k) {
            size_t i = IX(M, j, k), i1 = IX(M, jc, k);
            double sc = qj[i1] / qj[i];
            for (int n = 0; n < 4; ++n) ss[n][i] = ss[n][i1] * sc;
          }
        }
      } else {
        for (int k = ks; k <= ke; ++k) {
          int kc = periodic_src(idir, js, je, ks, ke, M->jd, M->kd, k);
          for (int j = js; j <= je; ++j) {
            size_t i = IX(M, j, k), i1 = IX(M, j, kc);
            double sc = qj[i1] / qj[i];
            for (int n = 0; n < 4; ++n) ss[n][i] = ss[n][i1] * sc;
          }
        }
      }
    } else if (t == 51 && iadir == 2) {
      for (int kc = 1; kc <= ke - ks + 1; ++kc) {
        int k, k1;
        if (idir == 2) { k = ke - kc + 1; k1 = ke + kc; } else { k = ks + kc - 1; k1 = ks - kc; }
        for (int j = js; j <= je; ++j) {
          int jj = M->jd - j + 1;
          size_t a = IX(M, j, k), bb = IX(M, jj, k1), cc = IX(M, jj, k), d = IX(M, j, k1);
          double scale1 = qj[bb] / qj[a];
          double scale2 = qj[d] / qj[cc];
          for (int n = 0; n < 4; ++n) ss[n][a] = ss[n][bb] * scale1;
          for (int n = 0; n < 4; ++n) ss[n][cc] = ss[n][d] * scale2;
        }
      }
    }
  }
}

int32_t or_rbc(or_case* c) {
  for (int im = 0; im < c->nmesh; ++im)
    for (int nsp = 0; nsp < c->nspec; ++nsp) rbc_mesh(c, &c->m[im], nsp);
  return OT2D_OK;
}

/* ----------------------------------------------------------- LU-SGS (a11) */
/* ilu2d, src/lhs.f:470-814 (iunst != 2 branch) */
static void ilu2d(or_case* c, or_mesh* M, int nsp) {
  int jd = M->jd, kd = M->kd, js = M->jbeg, je = M->jend, ks = M->kbeg, ke = M->kend;
  double gm1 = c->gm1, gamma = c->p.gamma, ggm1 = c->ggm1;
  double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
  double* qj = QA(M, nsp, c->nq);
  double* s[4] = {SA(M, nsp, 1), SA(M, nsp, 2), SA(M, nsp, 3), SA(M, nsp, 4)};
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  double *ug = GA(M, ug, nsp), *vg = GA(M, vg, nsp), *turmu = GA(M, turmu, nsp), *ts = GA(M, tscale, nsp);
  size_t n2 = M->n2;
  double* d = (double*)calloc(n2, sizeof(double));
  double* st = (double*)calloc(n2 * 4, sizeof(double));
  double* uv = (double*)calloc(n2, sizeof(double));
  double* vn = (double*)calloc(n2, sizeof(double));
  double* ge = (double*)calloc(n2, sizeof(double));
  double* qx = (double*)calloc(n2, sizeof(double));
  double* qy = (double*)calloc(n2, sizeof(double));
  double* cx = (double*)calloc(n2, sizeof(double));
  double* cy = (double*)calloc(n2, sizeof(double));
  int mdim = c->mdim + 2;
  double* a = (double*)calloc((size_t)mdim * 4, sizeof(double));
  double* b = (double*)calloc((size_t)mdim * 4, sizeof(double));
  double eps2 = c->p.epse * 2.;
  double epsv = 1. + eps2;
  double rmue = c->rmue, rey = c->rey;
  (void)jd; (void)kd;
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      double dj = 1.0 / q1[i];
      q2[i] = q2[i] * dj; q3[i] = q3[i] * dj; q4[i] = q4[i] * dj;
      q1[i] = q1[i] * qj[i];
    }
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      double uu = q2[i], vv = q3[i];
      double uv2 = 0.5 * (uu * uu + vv * vv);
      double cjkl = sqrt(ggm1 * (q4[i] - uv2));
      double rj1 = xx[i], rj2 = xy[i];
      double qq1 = rj1 * uu + rj2 * vv;
      double qqx = fabs(-rj1 * ug[i] - rj2 * vg[i] + qq1);
      double rr1 = sqrt(sq(rj1) + sq(rj2));
      double rk1 = yx[i], rk2 = yy[i];
      double qq2 = rk1 * uu + rk2 * vv;
      double qqy = fabs(-rk1 * ug[i] - rk2 * vg[i] + qq2);
      double rr2 = sqrt(sq(rk1) + sq(rk2));
      double vnu = 2.0 * rr2 * rr2 * (rmue + turmu[i]) / (rey * q1[i]);
      double svt = ts[i];
      d[i] = 1.0 / (1.0 + svt * ((qqx + qqy + cjkl * (rr1 + rr2)) * epsv + vnu));
      uv[i] = uv2; qx[i] = qq1; cx[i] = cjkl * rr1; qy[i] = qq2; cy[i] = cjkl * rr2; vn[i] = vnu;
      ge[i] = gamma * q4[i] - gm1 * uv2;
    }
  for (int k = ks; k <= ke; ++k) {
    size_t i = IX(M, js - 1, k), i2 = IX(M, je + 1, k);
    for (int n = 0; n < 4; ++n) s[n][i] = s[n][i] * d[i];
    for (int n = 0; n < 4; ++n) s[n][i2] = s[n][i2] * d[i2];
  }
  for (int j = js; j <= je; ++j) {
    size_t i = IX(M, j, ks - 1), i2 = IX(M, j, ke + 1);
    for (int n = 0; n < 4; ++n) s[n][i] = s[n][i] * d[i];
    for (int n = 0; n < 4; ++n) s[n][i2] = s[n][i2] * d[i2];
  }
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      for (int n = 0; n < 4; ++n) st[n * n2 + i] = s[n][i];
    }
  /* forward sweep over hyperplanes j+k = m */
  for (int m = js + ks; m <= je + ke; ++m) {
    int ms = MAX(js, m - ke), me = MIN(je, m - js);
    for (int j = ms; j <= me; ++j) {
      int k = m - j;
      size_t i1 = IX(M, j - 1, k);
      double uu = q2[i1], vv = q3[i1], uv2 = uv[i1], ri1 = xx[i1], ri2 = xy[i1], qq = qx[i1], cc = cx[i1];
      double qqx = -ri1 * ug[i1] - ri2 * vg[i1] + qq;
      double sp1 = fabs(qqx) + cc;
      double sm2 = eps2 * sp1;
      double chkx = 0.5 + fsign(0.5, qqx + cc);
      double spec = chkx * (qqx + sp1) + sm2;
      double s1 = s[0][i1], s2 = s[1][i1], s3 = s[2][i1], s4 = s[3][i1];
      double a4 = chkx * (ri1 * s2 + ri2 * s3 - qq * s1);
      double a2 = chkx * gm1 * (uv2 * s1 - (uu * s2 + vv * s3) + s4);
      a[j] = a4 + spec * s1;
      a[mdim + j] = ri1 * a2 + uu * a4 + spec * s2;
      a[2 * mdim + j] = ri2 * a2 + vv * a4 + spec * s3;
      a[3 * mdim + j] = qq * a2 + ge[i1] * a4 + spec * s4;
    }
    for (int j = ms; j <= me; ++j) {
      int k = m - j;
      size_t i1 = IX(M, j, k - 1);
      double uu = q2[i1], vv = q3[i1], uv2 = uv[i1], ri1 = yx[i1], ri2 = yy[i1], qq = qy[i1], cc = cy[i1];
      double vnu = vn[i1];
      double qqy = -ri1 * ug[i1] - ri2 * vg[i1] + qq;
      double sp1 = fabs(qqy) + cc;
      double sm2 = eps2 * sp1;
      double chky = 0.5 + fsign(0.5, qqy + cc);
      double spec = chky * (qqy + sp1) + vnu + sm2;
      double s1 = s[0][i1], s2 = s[1][i1], s3 = s[2][i1], s4 = s[3][i1];
      double a4 = chky * (ri1 * s2 + ri2 * s3 - qq * s1);
      double a2 = chky * gm1 * (uv2 * s1 - (uu * s2 + vv * s3) + s4);
      b[j] = a4 + spec * s1;
      b[mdim + j] = ri1 * a2 + uu * a4 + spec * s2;
      b[2 * mdim + j] = ri2 * a2 + vv * a4 + spec * s3;
      b[3 * mdim + j] = qq * a2 + ge[i1] * a4 + spec * s4;
    }
    for (int j = ms; j <= me; ++j) {
      int k = m - j;
      size_t i = IX(M, j, k);
      double sv = ts[i] * 0.5;
      double di = d[i];
      for (int n = 0; n < 4; ++n)
        s[n][i] = (st[n * n2 + i] + sv * (a[n * mdim + j] + b[n * mdim + j])) * di;
    }
  }
  /* backward sweep */
  for (int m = je + ke; m >= js + ks; --m) {
    int ms = MAX(js, m - ke), me = MIN(je, m - js);
    for (int j = ms; j <= me; ++j) {
      int k = m - j;
      size_t i1 = IX(M, j + 1, k);
      double uu = q2[i1], vv = q3[i1], uv2 = uv[i1], ri1 = xx[i1], ri2 = xy[i1], qq = qx[i1], cc = cx[i1];
      double qqx = -ri1 * ug[i1] - ri2 * vg[i1] + qq;
      double sp1 = fabs(qqx) + cc;
      double sm2 = eps2 * sp1;
      double chkx = 0.5 - fsign(0.5, qqx - cc);
      double spec = chkx * (qqx - sp1) - sm2;
      double s1 = s[0][i1], s2 = s[1][i1], s3 = s[2][i1], s4 = s[3][i1];
      double a4 = chkx * (ri1 * s2 + ri2 * s3 - qq * s1);
      double a2 = chkx * gm1 * (uv2 * s1 - (uu * s2 + vv * s3) + s4);
      a[j] = a4 + spec * s1;
      a[mdim + j] = ri1 * a2 + uu * a4 + spec * s2;
      a[2 * mdim + j] = ri2 * a2 + vv * a4 + spec * s3;
      a[3 * mdim + j] = qq * a2 + ge[i1] * a4 + spec * s4;
    }
    for (int j = ms; j <= me; ++j) {
      int k = m - j;
      size_t i1 = IX(M, j, k + 1);
      double uu = q2[i1], vv = q3[i1], uv2 = uv[i1], ri1 = yx[i1], ri2 = yy[i1], qq = qy[i1], cc = cy[i1];
      double vnu = vn[i1];
      double qqy = -ri1 * ug[i1] - ri2 * vg[i1] + qq;
      double sp1 = fabs(qqy) + cc;
      double sm2 = eps2 * sp1;
      double chky = 0.5 - fsign(0.5, qqy - cc);
      double spec = chky * (qqy - sp1) - vnu - sm2;
      double s1 = s[0][i1], s2 = s[1][i1], s3 = s[2][i1], s4 = s[3][i1];
      double a4 = chky * (ri1 * s2 + ri2 * s3 - qq * s1);
      double a2 = chky * gm1 * (uv2 * s1 - (uu * s2 + vv * s3) + s4);
      b[j] = a4 + spec * s1;
      b[mdim + j] = ri1 * a2 + uu * a4 + spec * s2;
      b[2 * mdim + j] = ri2 * a2 + vv * a4 + spec * s3;
      b[3 * mdim + j] = qq * a2 + ge[i1] * a4 + spec * s4;
    }
    for (int j = ms; j <= me; ++j) {
      int k = m - j;
      size_t i = IX(M, j, k);
      double sv = ts[i] * 0.5;
      double di = d[i];
      for (int n = 0; n < 4; ++n) s[n][i] = s[n][i] - sv * (a[n * mdim + j] + b[n * mdim + j]) * di;
    }
  }
  for (int k = ks - 1; k <= ke + 1; ++k)
    for (int j = js - 1; j <= je + 1; ++j) {
      size_t i = IX(M, j, k);
      q1[i] = q1[i] / qj[i];
      q2[i] = q2[i] * q1[i]; q3[i] = q3[i] * q1[i]; q4[i] = q4[i] * q1[i];
    }
  free(d); free(st); free(uv); free(vn); free(ge); free(qx); free(qy); free(cx); free(cy);
  free(a); free(b);
}

/* s *= tscale then ilu2d, src/step.f:141-159 */
static void lhs_mesh(or_case* c, or_mesh* M, int nsp) {
  double* ts = GA(M, tscale, nsp);
  for (int n = 1; n <= 4; ++n) {
    double* s = SA(M, nsp, n);
    for (size_t i = 0; i < M->n2; ++i) s[i] = s[i] * ts[i];
  }
  ilu2d(c, M, nsp);
}

int32_t or_lhs(or_case* c) {
  if (c->p.ilhs != 1 || c->p.iprecon) return OT2D_ERR_UNSUPPORTED;
  for (int im = 0; im < c->nmesh; ++im)
    for (int nsp = 0; nsp < c->nspec; ++nsp) lhs_mesh(c, &c->m[im], nsp);
  return OT2D_OK;
}

/* ------------------------------------------------------- monitor (a15) */
/* monitor, src/step.f:332-392 */
int32_t or_monitor(or_case* c, int32_t* mstop) {
  int stop = 0;
  for (int im = 0; im < c->nmesh; ++im) {
    or_mesh* M = &c->m[im];
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
      for (size_t i = 0; i < M->n2; ++i) {
        double rho = q1[i];
        double qq = sq(q2[i]) + sq(q3[i]);
        double asq = c->ggm1 * (q4[i] - .5 * qq / rho) / rho;
        if (asq <= 0.002 || asq + 1 == asq) stop = 1;
      }
    }
  }
  if (mstop) *mstop = stop;
  return OT2D_OK;
}

/* ---------------------------------------------------------- step (a9) */
/* residual norms, src/step.f:53-115 */
static int norms_mesh(or_case* c, or_mesh* M, int nsp, double* out, int32_t* am) {
  const int32_t* ib = M->iblank + (size_t)nsp * M->n2;
  double rsum = 0., resrho = 0., resmax = 0.;
  int jd1 = 2, kd1 = 2, nd1 = 1;
  for (int k = M->kbeg; k <= M->kend; ++k) {
    for (int n = 1; n <= c->nmv; ++n) {
      const double* s = SA(M, nsp, n);
      for (int j = M->jbeg; j <= M->jend; ++j) {
        size_t i = IX(M, j, k);
        double smrs = fabs(s[i]) * (double)MAX(ib[i], 0);
        if (smrs > resmax) { jd1 = j; kd1 = k; nd1 = n; resmax = smrs; }
      }
    }
    const double *s1 = SA(M, nsp, 1), *s2 = SA(M, nsp, 2), *s3 = SA(M, nsp, 3), *s4 = SA(M, nsp, 4);
    for (int j = M->jbeg; j <= M->jend; ++j) {
      size_t i = IX(M, j, k);
      double ibp = (double)MAX(ib[i], 0);
      double m1 = s1[i] * ibp, m2 = s2[i] * ibp, m3 = s3[i] * ibp, m4 = s4[i] * ibp;
      resrho = resrho + sq(s1[i]);
      rsum = rsum + m1 * m1 + m2 * m2 + m3 * m3 + m4 * m4;
    }
  }
  double volum = (double)(M->jd * M->kd);
  resrho = sqrt(resrho / volum);
  rsum = sqrt(rsum / volum);
  if (out) { out[0] = rsum; out[1] = resrho; out[2] = resmax; }
  if (am) { am[0] = jd1; am[1] = kd1; am[2] = nd1; }
  return rsum > 1000.0 ? OT2D_NORM : OT2D_OK;
}

/* step, src/step.f:2-218, with istep0 advanced first
 * (src/Turns_overset.f:364) and monitor after (src/Turns_overset.f:495) */
int32_t or_step(or_case* c, double* norms, int32_t* argmax) {
  c->istep0 += 1;
  c->istep += 1;
  int status = OT2D_OK;
  for (int im = 0; im < c->nmesh; ++im) {
    or_mesh* M = &c->m[im];
    time_step_mesh(c, M);
    memset(M->s, 0, M->n2 * c->nspec * c->nv * sizeof(double));
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      bc_mesh(c, M, nsp);
      int rc = rhsup_mesh(c, M, nsp);
      if (rc) return rc;
      if (!c->p.invisc) visrhs_mesh(c, M, nsp);
      if (c->timespectral) timespectralrhs(c, M, nsp);
    }
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      size_t o = (size_t)im * c->nspec + nsp;
      int rc = norms_mesh(c, M, nsp, norms ? norms + 3 * o : NULL, argmax ? argmax + 3 * o : NULL);
      if (rc == OT2D_NORM) return OT2D_NORM;
    }
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      /* s = s + ss with ss = 0 (src/step.f:126-132) */
      for (int n = 1; n <= c->nv; ++n) {
        double* s = SA(M, nsp, n);
        for (int k = M->kbeg; k <= M->kend; ++k)
          for (int j = M->jbeg; j <= M->jend; ++j) s[IX(M, j, k)] = s[IX(M, j, k)] + 0.0;
      }
      rbc_mesh(c, M, nsp);
      if (c->p.ilhs != 1 || c->p.iprecon) return OT2D_ERR_UNSUPPORTED;
      lhs_mesh(c, M, nsp);
    }
    if (!c->lamin) {
      if ((c->iturb != 1 && c->iturb != 2) || c->p.itrans) return OT2D_ERR_UNSUPPORTED;
      for (int nsp = 0; nsp < c->nspec; ++nsp) {
        if (c->iturb == 1) vmu_sa_mesh(c, M, nsp);
        else {
          int rc = komegasst_mesh(c, M, nsp);
          if (rc) return rc;
        }
      }
    }
    for (int nsp = 0; nsp < c->nspec; ++nsp) {
      const int32_t* ib = M->iblank + (size_t)nsp * M->n2;
      for (int n = 1; n <= 4; ++n) {
        double* q = QA(M, nsp, n);
        const double* s = SA(M, nsp, n);
        for (int k = M->kbeg; k <= M->kend; ++k)
          for (int j = M->jbeg; j <= M->jend; ++j) {
            size_t i = IX(M, j, k);
            q[i] = q[i] + s[i] * (double)MAX(ib[i], 0);
          }
      }
    }
    int32_t ms = 0;
    /* monitor per mesh, as the driver calls it right after step */
    {
      int stop = 0;
      for (int nsp = 0; nsp < c->nspec; ++nsp) {
        double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
        for (size_t i = 0; i < M->n2; ++i) {
          double rho = q1[i];
          double qq = sq(q2[i]) + sq(q3[i]);
          double asq = c->ggm1 * (q4[i] - .5 * qq / rho) / rho;
          if (asq <= 0.002 || asq + 1 == asq) stop = 1;
        }
      }
      ms = stop;
    }
    if (ms) { status = OT2D_MSTOP; break; }
  }
  /* after every mesh has stepped (src/Turns_overset.f:500-503) */
  if (status == OT2D_OK && c->overset) do_interpolations(c);
  return status;
}

int32_t or_iterate(or_case* c, int32_t niter, double* hist, int32_t* done) {
  size_t per = (size_t)c->nmesh * c->nspec;
  double* nb = (double*)calloc(per * 3, sizeof(double));
  int rc = OT2D_OK, it;
  for (it = 0; it < niter; ++it) {
    rc = or_step(c, nb, NULL);
    if (hist)
      for (size_t o = 0; o < per; ++o) hist[(size_t)it * per + o] = nb[3 * o];
    if (rc != OT2D_OK) { ++it; break; }
  }
  if (done) *done = it;
  free(nb);
  return rc;
}

/* ------------------------------------------------------- forces (a16) */
/* compute_forces + force2d, src/forces.f:2-107, 111-398 */
int32_t or_forces(or_case* c, int32_t im, int32_t nsp, double out[6]) {
  if (!c || im < 0 || im >= c->nmesh || nsp < 0 || nsp >= c->nspec) return OT2D_ERR_ARG;
  or_mesh* M = &c->m[im];
  double *q1 = QA(M, nsp, 1), *q2 = QA(M, nsp, 2), *q3 = QA(M, nsp, 3), *q4 = QA(M, nsp, 4);
  double* qj = QA(M, nsp, c->nq);
  const double* xv = VA(M, xv, nsp);
  const double* yv = VA(M, yv, nsp);
  double *xx = GA(M, xx, nsp), *xy = GA(M, xy, nsp), *yx = GA(M, yx, nsp), *yy = GA(M, yy, nsp);
  double gm1 = c->gm1, gamma = c->p.gamma;
  int k = M->kbeg, k2 = k + 1, k3 = k2 + 1;
  int j1 = M->jtail1, j2 = M->jtail2;
  double fmm = c->p.fsmach != 0 ? c->p.fsmach : (c->p.fmtip != 0 ? c->p.fmtip : 1.0);
  double half = (double)c->p.half;
  double* zz = (double*)calloc((size_t)M->jd + 2, sizeof(double));
  double* sxx = (double*)calloc((size_t)M->jd + 2, sizeof(double));
  double* syy = (double*)calloc((size_t)M->jd + 2, sizeof(double));
#define XVv(j, k) xv[VX(M, j, k)]
#define YVv(j, k) yv[VX(M, j, k)]
  for (int j = j1; j <= j2; ++j) {
    size_t i = IX(M, j, k), im1 = IX(M, j, k - 1);
    double rr = 1. / q1[i];
    double rho = q1[i] * qj[i];
    double u = q2[i] * rr, v = q3[i] * rr, e = q4[i] * qj[i];
    double vsq = u * u + v * v;
    double pp = gm1 * (e - rho * vsq / 2.);
    double pp1 = pp / c->pinf;
    double cp = 2. * (pp1 - 1.) / (gamma * sq(fmm));
    zz[j] = cp;
    rr = 1. / q1[im1];
    rho = q1[im1] * qj[im1];
    u = q2[im1] * rr; v = q3[im1] * rr; e = q4[im1] * qj[im1];
    vsq = u * u + v * v;
    pp = gm1 * (e - rho * vsq / 2.);
    pp1 = pp / c->pinf;
    cp = 2. * (pp1 - 1.) / (gamma * sq(fmm));
    zz[j] = 0.5 * (zz[j] + cp);
  }
  double cn = 0., cc = 0., cmle = 0., cpinv = 0.;
  for (int j = j1; j <= j2; ++j) {
    int jv = j - 1, kv = k - 1;
    cn = cn - zz[j] * (XVv(jv + 1, kv) - XVv(jv, kv));
    cc = cc + zz[j] * (YVv(jv + 1, kv) - YVv(jv, kv));
    cmle = cmle + zz[j] * (0.5 * (XVv(jv + 1, kv) + XVv(jv, kv)) * (XVv(jv + 1, kv) - XVv(jv, kv)) +
                           0.5 * (YVv(jv + 1, kv) + YVv(jv, kv)) * (YVv(jv + 1, kv) - YVv(jv, kv)));
    size_t i = IX(M, j, k), im1 = IX(M, j, k - 1);
    double u = q2[i] / q1[i], v = q3[i] / q1[i];
    double um = q2[im1] / q1[im1], vm = q3[im1] / q1[im1];
    cpinv = cpinv + 0.5 * (u + um) * zz[j] * (YVv(jv + 1, kv) - YVv(jv, kv))
            - 0.5 * (v + vm) * zz[j] * (XVv(jv + 1, kv) - XVv(jv, kv));
  }
  cn = cn - half * cn;
  cmle = cmle - half * cmle;
  cc = cc + half * cc;
  double cnv = 0., ccv = 0., cmlev = 0., cpv = 0.;
  if (!c->p.invisc) {
    double alngth = 1.;
    double amu = c->rinf * alngth / c->rey;
    double uinf2 = sq(fmm);
#define UQ(j, kk) (q2[IX(M, j, kk)] / q1[IX(M, j, kk)])
#define VQ(j, kk) (q3[IX(M, j, kk)] / q1[IX(M, j, kk)])
    for (int j = j1; j <= j2; ++j) {
      double uxi, vxi;
      if (j == j1) {
        uxi = UQ(j + 1, k) - UQ(j, k);
        vxi = VQ(j + 1, k) - VQ(j, k);
      } else if (j == j2) {
        uxi = UQ(j, k) - UQ(j - 1, k);
        vxi = VQ(j, k) - VQ(j - 1, k);
      } else {
        uxi = .5 * (UQ(j + 1, k) - UQ(j - 1, k));
        vxi = .5 * (VQ(j + 1, k) - VQ(j - 1, k));
      }
      /* -1.5*q(j,k,2)/q(j,k,1): (-1.5*q2)/q1, src/forces.f:299-302 */
      size_t i1 = IX(M, j, k), i2 = IX(M, j, k2), i3 = IX(M, j, k3);
      double ueta = -1.5 * q2[i1] / q1[i1] + 2. * q2[i2] / q1[i2] - .5 * q2[i3] / q1[i3];
      double veta = -1.5 * q3[i1] / q1[i1] + 2. * q3[i2] / q1[i2] - .5 * q3[i3] / q1[i3];
      size_t i = IX(M, j, k);
      double xix = xx[i], xiy = xy[i], etax = yx[i], etay = yy[i];
      double sfdiv = 2. / 3 * (uxi * xix + ueta * etax + vxi * xiy + veta * etay);
      sxx[j] = amu * (2. * (uxi * xix + ueta * etax) - sfdiv);
      syy[j] = amu * (2. * (vxi * xiy + veta * etay) - sfdiv);
      double tauw = amu * ((uxi * xiy + ueta * etay) + (vxi * xix + veta * etax));
      sxx[j] = sxx[j] / (.5 * c->rinf * uinf2);
      syy[j] = syy[j] / (.5 * c->rinf * uinf2);
      zz[j] = tauw / (.5 * c->rinf * uinf2);
    }
    for (int j = j1; j <= j2; ++j) {
      int jv = j - 1, kv = k - 1;
      double dxv = XVv(jv + 1, kv) - XVv(jv, kv), dyv = YVv(jv + 1, kv) - YVv(jv, kv);
      ccv = ccv + zz[j] * (XVv(jv + 1, kv) - XVv(jv, kv)) - sxx[j] * (YVv(jv + 1, kv) - YVv(jv, kv));
      cnv = cnv - zz[j] * (YVv(jv + 1, kv) - YVv(jv, kv)) + syy[j] * (XVv(jv + 1, kv) - XVv(jv, kv));
      cmlev = cmlev + zz[j] * (0.5 * (XVv(jv + 1, kv) + XVv(jv, kv)) * (YVv(jv + 1, kv) - YVv(jv, kv)) +
                               0.5 * (YVv(jv + 1, kv) + YVv(jv, kv)) * (XVv(jv + 1, kv) - XVv(jv, kv)))
              - 0.5 * (XVv(jv + 1, kv) + XVv(jv, kv)) * (XVv(jv + 1, kv) - XVv(jv, kv)) * syy[j]
              - 0.5 * (YVv(jv + 1, kv) + YVv(jv, kv)) * (YVv(jv + 1, kv) - YVv(jv, kv)) * sxx[j];
      size_t i = IX(M, j, k), im1 = IX(M, j, k - 1);
      double u = q2[i] / q1[i], v = q3[i] / q1[i];
      double um = q2[im1] / q1[im1], vm = q3[im1] / q1[im1];
      cpv = cpv + 0.5 * (u + um) * zz[j] * dxv - 0.5 * (u + um) * sxx[j] * dyv
            - 0.5 * (v + vm) * zz[j] * dyv + 0.5 * (v + vm) * syy[j] * dxv;
    }
    cnv = cnv - half * cnv;
    cmlev = cmlev - half * cmlev;
    ccv = ccv + half * ccv;
#undef UQ
#undef VQ
  }
#undef XVv
#undef YVv
  double cfx = cc + ccv, cfy = cn + cnv, cm0 = cmle + cmlev;
  double cpower = (cpinv + cpv) / fmm;
  double ca = cos(c->pi * c->p.alfa / 180.0), sa = sin(c->pi * c->p.alfa / 180.0);
  double x0 = 0.25, y0 = 0.00;
  out[0] = -cfx * sa + cfy * ca;
  out[1] = cfx * ca + cfy * sa;
  out[2] = cm0 - y0 * cfx + x0 * cfy;
  out[3] = cpower;
  out[4] = cfx;
  out[5] = cfy;
  free(zz); free(sxx); free(syy);
  return OT2D_OK;
}
