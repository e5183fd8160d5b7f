/*
 * ot2d.h -- C-ABI of the MI355X-native OVERTURNS-2D iteration
 * (residual + LU-SGS implicit update, SA turbulence, forces), gfx950 HIP.
 *
 * The boundary replaces the F77 calls the reference driver makes inside its
 * time loop (src/Turns_overset.f:475-554):
 *   time_step(q,xx,xy,yx,yy,ug,vg,tscale,bt,iblank,jd,kd)      src/time.f:2
 *   step(q,...,s,...,turmu,tscale,bt,Ds,im,jd,kd)               src/step.f:2-7
 *   monitor(mstop,q,jd,kd)                                      src/step.f:332
 *   compute_forces(jd,kd,x,y,xv,yv,q,xx,...,cl,cd,...)          src/forces.f:2-3
 * and the set-up it does once (initia, src/init.f:2-177; metfv,
 * src/setup.f:440; find_cellcenter, src/setup.f:111; qzero/qdivj,
 * src/init.f:180-277).
 *
 * Conventions (those of the reference, so host arrays can be passed as is):
 *  - arrays are Fortran column-major, cell-centred (jd,kd) with jd=jmax+1,
 *    kd=kmax+1 (one halo ring, nhalo=1), variable index slowest:
 *    q(jd,kd,nspec,nq), s(jd,kd,nspec,nv); vertex arrays (jmax,kmax,nspec);
 *  - q(:,:,:,1:4) hold conserved variables times cell area, q(:,:,:,nq) is
 *    1/area (src/init.f:197-203, src/setup.f:489-500); q(:,:,:,5..) are the
 *    turbulence variables as the reference stores them;
 *  - reals are IEEE binary64, integers int32_t, logicals int32_t 0/1;
 *  - every entry point returns an int32_t status: OT2D_OK, or
 *    OT2D_MSTOP (negative speed of sound, maps to the reference's mstop=1,
 *    src/step.f:353-385), OT2D_NORM (l2 norm > 1000, the reference's
 *    "stop 'norm'", src/step.f:105-110), or a negative error code.
 *  - a case handle is bound to the HIP device current at creation and to one
 *    HIP stream; calls on one handle are not re-entrant.
 */
#ifndef OT2D_H
#define OT2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OT2D_MAXBC 25

enum {
  OT2D_OK = 0,
  OT2D_MSTOP = 1,
  OT2D_NORM = 2,
  OT2D_CONVERGED = 3,   /* batch: case retired (ot2d_batch_retire) */
  OT2D_ERR_ARG = -1,
  OT2D_ERR_HIP = -2,
  OT2D_ERR_UNSUPPORTED = -3,
  OT2D_ERR_STATE = -4,
  OT2D_ERR_NOMEM = -5
};

/* One &bcinp row (src/setup.f:894-931; index rules src/bc.f:21-35). */
typedef struct ot2d_bcseg {
  int32_t ibtyp, ibdir, jbcs, jbce, kbcs, kbce;
} ot2d_bcseg;

/* Per-mesh description: vertex dimensions and the bc.inp table. */
typedef struct ot2d_mesh_desc {
  int32_t jmax, kmax;
  int32_t nbc;
  int32_t irot_dir;
  double xcen, ycen;
  ot2d_bcseg bc[OT2D_MAXBC];
} ot2d_mesh_desc;

/* The &inputs namelist values the iteration reads (src/modules.f:170-208)
 * plus the fixed data of params_global.  Fill with ot2d_default_params()
 * (the reference's data statements) and override what the deck sets. */
typedef struct ot2d_params {
  int32_t iread, nhalo, half;
  int32_t invisc, lamin, iturb, itrans;
  int32_t irhsy, ilim, ilhs, iconstant_cfl, idual, iprecon;
  int32_t ntac, iunst, itnmax, timespectral, nspec, motiontype;
  int32_t nsteps, npnorm, nrest;
  double fsmach, alfa, rey, vnuinf, tuinf;
  double dt, timeac, epse, dualtime, Mp;
  double rf, angmax, fmtip, rartio, totime;
  double gamma, pr, rmue, tinf;
  /* extensions (no reference counterpart): arithmetic of the LU-SGS
   * operator, 0 = binary64 (reference), 1 = binary32 sweeps */
  int32_t lu_fp32;
  /* static_conn of &inputs (src/modules.f:186, 204): 1 keeps the first
   * connectivity of a moving overset case instead of cutting again after
   * every step's motion (src/Turns_overset.f:413) */
  int32_t static_conn;
  int32_t reserved[6];
} ot2d_params;

typedef struct ot2d_case ot2d_case;

/* Fields readable with ot2d_get_field (all (jd,kd,nspec) unless noted). */
enum {
  OT2D_F_X = 0, OT2D_F_Y, OT2D_F_XX, OT2D_F_XY, OT2D_F_YX, OT2D_F_YY,
  OT2D_F_UG, OT2D_F_VG, OT2D_F_TSCALE, OT2D_F_BT, OT2D_F_TURMU,
  OT2D_F_S,          /* (jd,kd,nspec,nv) */
  OT2D_F_XV, OT2D_F_YV, /* (jmax,kmax,nspec) */
  OT2D_F_UGV, OT2D_F_VGV,
  OT2D_F_COUNT
};

/* Defaults of params_global data statements (src/modules.f:170-193). */
int32_t ot2d_default_params(ot2d_params* p);

/* Derived sizes for mesh im of a case: jd, kd, nq, nv. */
int32_t ot2d_sizes(const ot2d_params* p, const ot2d_mesh_desc* m,
                   int32_t* jd, int32_t* kd, int32_t* nq, int32_t* nv);

/* Allocate device state for nmesh meshes (replaces the allocation block,
 * src/Turns_overset.f:191-237, and read_inputs' derived constants,
 * src/setup.f:1162-1265). */
int32_t ot2d_case_create(const ot2d_params* p, int32_t nmesh,
                         const ot2d_mesh_desc* meshes, ot2d_case** out);
int32_t ot2d_case_destroy(ot2d_case* c);

/* Upload the vertex grid of mesh im (jmax,kmax) and initialise it as
 * initia does (src/init.f:2-177): cell centres, finite-volume metrics,
 * free stream (qzero), q/J (qdivj), iblank=1, turmu=ug=vg=0, istep0=0. */
int32_t ot2d_set_grid(ot2d_case* c, int32_t im, const double* xv, const double* yv);

/* Internal state q(jd,kd,nspec,nq) of mesh im (as the reference holds it). */
int32_t ot2d_set_state(ot2d_case* c, int32_t im, const double* q);
int32_t ot2d_get_state(ot2d_case* c, int32_t im, double* q);
int32_t ot2d_set_iblank(ot2d_case* c, int32_t im, const int32_t* iblank);
int32_t ot2d_get_iblank(ot2d_case* c, int32_t im, int32_t* iblank);
int32_t ot2d_get_field(ot2d_case* c, int32_t im, int32_t field, double* out);
int32_t ot2d_set_field(ot2d_case* c, int32_t im, int32_t field, const double* in);

/* istep0 counter (src/Turns_overset.f:364; drives the wall-BC ramp,
 * src/bc.f:1779-1793). */
int32_t ot2d_set_istep0(ot2d_case* c, int32_t istep0);
int32_t ot2d_get_istep0(ot2d_case* c, int32_t* istep0);

/* Stage entry points, one per reference subroutine, acting on the
 * device-resident state of every mesh of the case (used by parity tests
 * and by drivers that interleave their own work). */
int32_t ot2d_time_step(ot2d_case* c);                     /* src/time.f:2       */
int32_t ot2d_bc(ot2d_case* c);                            /* src/bc.f:2         */
int32_t ot2d_compute_residual(ot2d_case* c);              /* src/step.f:221     */
int32_t ot2d_rbc(ot2d_case* c);                           /* src/bc.f:88        */
int32_t ot2d_lhs(ot2d_case* c);    /* s*=tscale, ilu2d   src/step.f:141-159, src/lhs.f:470 */
int32_t ot2d_turbulence(ot2d_case* c);  /* solve_turbulence_equations, src/step.f:303 (SA) */
int32_t ot2d_monitor(ot2d_case* c, int32_t* mstop);       /* src/step.f:332     */

/* One full step of every mesh (src/step.f:2-218) followed by monitor, as
 * the driver's inner loop does (src/Turns_overset.f:475-498).  norms
 * receives rsum,resrho,resmax per mesh and instance (3*nmesh*nspec
 * doubles, may be NULL); argmax receives j,k,n of resmax (3 ints each, may
 * be NULL).  istep0 is advanced by one first. */
int32_t ot2d_step(ot2d_case* c, double* norms, int32_t* argmax);

/* Time-accurate runs (itnmax > 1 Newton sub-iterations, ntac = 2, 3 for
 * BDF2 / BDF3 time accuracy, or a pitching grid, iunst = 1 with
 * motiontype = 1; not time-spectral): every step of ot2d_step /
 * ot2d_iterate is one time step of the reference's driver loop
 * (src/Turns_overset.f:357-510): a pitching grid turns by the step's angle
 * (pitch, src/motion.f:546-632) and its metrics, grid velocities and wall
 * distances follow, the time levels shift (stqol, src/time.f:109-172), then
 * itnmax sub-iterations of time_step, step (with the Newton time term,
 * src/time.f:174-209, and the swept-area face speeds of the moving grid,
 * src/eulerrhs.f:118-136) and monitor.  The norms ot2d_step returns are the last sub-iteration's.
 * A pitching overset case (several meshes, after ot2d_connect) moves its
 * first mesh only, and ot2d_step / ot2d_iterate find its connectivity again
 * on the host after every step's motion, with the ihcbc rows of the last
 * ot2d_connect (src/Turns_overset.f:411-465); ot2d_iterate_async,
 * ot2d_iterate_noalloc and ot2d_profile refuse such a case
 * (OT2D_ERR_UNSUPPORTED).
 * ot2d_set_dtpseudo sets the dual time of sub-iterations 1..n
 * (dtpseudo(itn) of the deck; default: dualtime for all, n <= 16). */
int32_t ot2d_set_dtpseudo(ot2d_case* c, int32_t n, const double* dtpseudo);
/* Norms of every sub-iteration of the last time step:
 * norms[(itn*nmesh*nspec + i)*3 + m] (rsum, resrho, resmax) and argmax in
 * the same order (j, k, n), itn = 0..itnmax-1.  Either may be NULL. */
int32_t ot2d_subiter_norms(ot2d_case* c, double* norms, int32_t* argmax);

/* niter steps without host synchronisation in between; the residual
 * history (rsum per step per mesh/instance) is written to hist
 * (niter*nmesh*nspec doubles, may be NULL).  Stops early on MSTOP/NORM and
 * reports the number of steps done in *done. */
int32_t ot2d_iterate(ot2d_case* c, int32_t niter, double* hist, int32_t* done);

/* Sectional forces of mesh im, instance nsp (0-based): out = cl, cd, cm,
 * cpower, cfx, cfy (src/forces.f:2-107, force2d src/forces.f:111-398). */
int32_t ot2d_forces(ot2d_case* c, int32_t im, int32_t nsp, double out[6]);

/* Batches: ncase independent single-mesh cases advanced together by the
 * same kernels (one launch per stage for all of them; a case of uneven size
 * only occupies its own tiles, lines and sweep workgroups).  Each case has
 * its own flow conditions, grid, step counter and status; the cases must
 * share the model, the LU-SGS precision and the instance count.  Case i is
 * mesh i for every per-mesh call (ot2d_set_grid, ot2d_get_state,
 * ot2d_forces, ...).  ot2d_iterate advances every running case and writes
 * rsum per step and case to hist; a case that stops (MSTOP, NORM) or is
 * retired stays so until its grid or state is set again.  No counterpart
 * in the reference, which runs one case per process. */
int32_t ot2d_batch_create(int32_t ncase, const ot2d_params* params, const ot2d_mesh_desc* meshes,
                          ot2d_case** out);
/* Per case: OT2D_OK (running), OT2D_MSTOP, OT2D_NORM, OT2D_CONVERGED, or a
 * negative error; istep0 of each case.  Either pointer may be NULL. */
int32_t ot2d_batch_status(ot2d_case* c, int32_t* status, int32_t* istep0);
/* Stop advancing case icase (status OT2D_CONVERGED). */
int32_t ot2d_batch_retire(ot2d_case* c, int32_t icase);
/* Launch the batch's following steps over its running cases only: waits
 * for the case's stream, reads every case's status, and keeps the running
 * cases' instance records and residual tiles, so stopped cases (converged,
 * retired, failed) cost no workgroups.  Calls that may make a stopped case
 * run again (ot2d_set_grid, ot2d_set_state, ot2d_batch_admit) return the
 * batch to launching over every case.  *running (may be NULL): cases
 * running.  Batches without overset connectivity.  No counterpart in the
 * reference; for request loops whose batches hold finished cases. */
int32_t ot2d_batch_compact(ot2d_case* c, int32_t* running);
/* Put a new case into slot icase of a batch, whose case has stopped
 * (converged, retired or failed): parameters p (same model, precision and
 * instance count as the batch) and mesh md, which must fit the memory the
 * slot holds (its cells, vertices, LU workspace and residual tiles, and no
 * more rows per lane of the FP64 sweep than the batch's).  The slot's
 * memory is cleared as for a new case, its step counter and status reset,
 * and the captured step re-made at the next iteration; call ot2d_set_grid
 * (and ot2d_set_convergence) for it next.  Batches of one mesh per case
 * without overset connectivity.  For a request loop whose new requests take
 * the slots of finished ones; no counterpart in the reference. */
int32_t ot2d_batch_admit(ot2d_case* c, int32_t icase, const ot2d_params* p, const ot2d_mesh_desc* md);

/* ------------------------------------------------------------ launch choices
 * Launch shapes that change speed but not results: sweep (0, 1 or 2) picks
 * the FP32 LU-SGS sweep's LDS ring depth and producer count (12 slots and
 * 3 producer waves, 18 and 3, 12 and 4); sa_lane picks the SA line solves'
 * mapping (-1 by line count, 0 one workgroup per line, 1 one lane per line).
 * Every choice gives bit-identical states.  No counterpart in the reference,
 * whose loops have one shape. */
int32_t ot2d_set_launch(ot2d_case* c, int32_t sweep, int32_t sa_lane);
int32_t ot2d_get_launch(ot2d_case* c, int32_t* sweep, int32_t* sa_lane);
/* Measured search over the launch choices for the case's shape class (model,
 * precision, instances, case count, and cells, jmax and sweep blocks):
 * each candidate times `steps` captured steps from the case's current state
 * (HIP events on the case's stream, best of three interleaved rounds), the
 * state (q, eddy viscosity, step counters, status) is put back after each,
 * and the fastest is set on the case.  Results are cached per class in
 * the process and, when OT2D_TUNE_CACHE names a file, in that file (one line
 * per class, read at the first call, appended after each search); a cached
 * class is set without measuring.  Either way the chosen step is captured
 * before the call returns, so the next ot2d_iterate starts at once.
 * *measured (may be NULL) is 1 when a search ran, 0 on a cache hit.
 * Allocates a copy of the state while it runs. */
int32_t ot2d_autotune(ot2d_case* c, int32_t steps, int32_t* measured);

/* --------------------------------------------------------------- serving
 * Primitives of a request loop that keeps many cases resident and admits and
 * retires them between iterations.  No counterpart in the reference (its
 * driver runs one case to nsteps); the convergence test is the reference's
 * residual norm rsum (src/step.f:60-93). */

/* Device bytes ot2d_case_create (ncase = 1) or ot2d_batch_create
 * (ncase = nmesh) would hold for these meshes, for admission by footprint. */
int32_t ot2d_footprint(int32_t ncase, const ot2d_params* params, int32_t nmesh, const ot2d_mesh_desc* meshes,
                       int64_t* bytes);
/* Device bytes the case holds now. */
int32_t ot2d_case_bytes(ot2d_case* c, int64_t* bytes);
/* Case icase (-1: all) stops, with status OT2D_CONVERGED, at the first step
 * whose rsum is at or below rsum_tol on every instance, before that step's
 * update; rsum_tol <= 0 switches the test off.  While the test is on, a
 * case whose rsum is not a finite number stops with OT2D_NORM. */
int32_t ot2d_set_convergence(ot2d_case* c, int32_t icase, double rsum_tol);
/* Enqueues niter steps on the case's stream and returns without waiting
 * (no residual history); ot2d_sync or ot2d_batch_status waits.  Cases on
 * different streams overlap on the device, except that the steps of cases
 * with lu_fp32 = 1 are ordered across such cases on one device (each
 * waits for the FP32 work enqueued before it by any case). */
int32_t ot2d_iterate_async(ot2d_case* c, int32_t niter);
int32_t ot2d_sync(ot2d_case* c);
/* *idle = 1 when everything enqueued for the case has finished, else 0;
 * never waits.  For request loops that advance each batch as soon as its
 * previous steps are done; no counterpart in the reference. */
int32_t ot2d_query(ot2d_case* c, int32_t* idle);
/* on = 1 (the default): the order of FP32 work across cases above; on = 0:
 * lifted for the whole process, so FP32 cases on different streams overlap
 * like FP64 ones.  Applies to work enqueued after the call.  No counterpart
 * in the reference. */
int32_t ot2d_set_fp32_order(int32_t on);

/* ot2d_iterate without allocating: the same steps, the same outputs (hist,
 * *done, status), but the device-side residual history is staged in a
 * fixed scratch array inside the library instead of a buffer the call
 * allocates, and the call makes no hipMalloc, hipFree, graph capture or
 * other allocating call.  No counterpart in the reference; for callers that
 * cannot allocate (tight serving loops, harnesses that forbid allocation).
 * Limit: niter * nmesh * nspec <= OT2D_NOALLOC_HIST_MAX when hist is not
 * NULL (OT2D_ERR_ARG otherwise).  The scratch is one per GPU, shared by
 * every case on that GPU: calls that overlap on one GPU race. */
#define OT2D_NOALLOC_HIST_MAX (1 << 20)
int32_t ot2d_iterate_noalloc(ot2d_case* c, int32_t niter, double* hist, int32_t* done);

/* Kernel stages of one step, in launch order (ot2d_profile). */
enum {
  OT2D_ST_BEGIN = 0,     /* step counter / status reset                        */
  OT2D_ST_TIME_STEP,     /* time_step                 src/time.f:2-107         */
  OT2D_ST_BC,            /* bc                        src/bc.f:2-85            */
  OT2D_ST_RESIDUAL,      /* rhsup (+ norms if inviscid) src/eulerrhs.f:2-280   */
  OT2D_ST_VISRHS,        /* visrhs                    src/visrhs.f:2-287       */
  OT2D_ST_TSRHS,         /* timespectralrhs           src/sources.f:2-27       */
  OT2D_ST_NORMS,         /* residual norms            src/step.f:53-115        */
  OT2D_ST_RBC,           /* rbc                       src/bc.f:88-177          */
  OT2D_ST_LU_PREP,       /* s*=tscale, ilu2d set-up   src/lhs.f:470-612        */
  OT2D_ST_LU_FORWARD,    /* ilu2d forward sweep       src/lhs.f:613-715        */
  OT2D_ST_LU_BACKWARD,   /* ilu2d backward sweep      src/lhs.f:716-802        */
  OT2D_ST_LU_FINISH,     /* ilu2d restore src/lhs.f:803-814 (+update, monitor if laminar) */
  OT2D_ST_SA_PRIM,       /* lamvis, u, v              src/samodel2d.f:61-72, 200-238 */
  OT2D_ST_SA_BC,         /* turbc                     src/samodel2d.f:827-1303 */
  OT2D_ST_SA_RHS,        /* vortic, convec, diffus, source src/samodel2d.f:101-131, 402-824 */
  OT2D_ST_SA_SOLVE_J,    /* lsolvej                   src/samodel2d.f:1306-1344 */
  OT2D_ST_SA_SOLVE_K,    /* lsolvek + clip            src/samodel2d.f:157-188, 1347-1385 */
  OT2D_ST_SA_EDDY,       /* turbc + c_turm            src/samodel2d.f:372-399  */
  OT2D_ST_UPDATE,        /* q += s                    src/step.f:205-211       */
  OT2D_ST_MONITOR,       /* monitor                   src/step.f:332-392       */
  OT2D_ST_INTERP,        /* do_interpolations         src/connectivity.f:2-189 */
  OT2D_NSTAGE
};

/* Device time of each kernel stage, measured with HIP events recorded on the
 * case's stream between the kernels of nsteps eagerly launched steps (the
 * same kernels ot2d_iterate replays; the case advances by those steps).
 * stage_ms[OT2D_NSTAGE] receives the mean milliseconds per step of every
 * stage (0 for stages this case does not run); *done = steps completed.
 * Instrumentation for benchmarks; no counterpart in the reference. */
int32_t ot2d_profile(ot2d_case* c, int32_t nsteps, double* stage_ms, int32_t* done);
const char* ot2d_stage_name(int32_t stage);

/* ------------------------------------------------------------ overset meshes
 * Connectivity of overlapping meshes by implicit hole cutting (connect2d,
 * ihc/src/ihc.f90:252-448, called from src/Turns_overset.f:294-353) and the
 * per-iteration donor interpolation (do_interpolations,
 * src/connectivity.f:2-189).  The hole cutting runs once per static overset
 * configuration on the host; the interpolation runs on the device inside
 * every step of a case whose connectivity is set. */

/* One ihcbc.<n> file (ihc/src/ihc.f90:147-208).  Types: 1 not a receiver,
 * 2 blanked, 3 forced receiver, 4 never blanked, 5 wall points (indices count
 * vertices), 6 = 1 and 4, 7 = 3 and 4.  Other ranges count cells without the
 * halo; negative values count from the end. */
typedef struct ot2d_ihcbc {
  int32_t nbc;
  int32_t ibtyp[OT2D_MAXBC], ibdir[OT2D_MAXBC];
  int32_t jbcs[OT2D_MAXBC], jbce[OT2D_MAXBC], kbcs[OT2D_MAXBC], kbce[OT2D_MAXBC];
} ot2d_ihcbc;

/* One mesh as connect2d reads it (src/Turns_overset.f:302-322): vertices
 * (jmax x kmax), cell centres and cell areas (1/J) of all (jmax+1) x (kmax+1)
 * cells including the halo ring, j fastest. */
typedef struct ot2d_ihc_mesh {
  int32_t jmax, kmax;
  const double *xv, *yv;
  const double *xc, *yc, *vol;
  ot2d_ihcbc bc;
} ot2d_ihc_mesh;

typedef struct ot2d_ihc_lists ot2d_ihc_lists;

/* connect2d over nmesh meshes on the host (the reference's module state is
 * rebuilt on every call).  *out owns the result; free with ot2d_ihc_free. */
int32_t ot2d_ihc_connect(const ot2d_ihc_mesh* meshes, int32_t nmesh, ot2d_ihc_lists** out);
/* Counts of mesh im (0-based): fringe cells, donor cells and the 1-based
 * range of its donors in the concatenated donor list (iisptr, iieptr). */
int32_t ot2d_ihc_counts(const ot2d_ihc_lists* l, int32_t im, int32_t* nfringe, int32_t* ndonor, int32_t* iisptr,
                        int32_t* iieptr);
/* Lists of mesh im in the reference's layout (src/Turns_overset.f:223-226):
 * iblank (jmax+1)*(kmax+1) after update_halo_iblanks (src/connectivity.f:192-234);
 * imesh (nfringe x 2) and idonor (ndonor x 2) 1-based solver cell indices,
 * column-major; ibc (nfringe) 1-based rows of the concatenated donor list;
 * frac (ndonor x 2).  Any pointer may be NULL. */
int32_t ot2d_ihc_get(const ot2d_ihc_lists* l, int32_t im, int32_t* iblank, int32_t* imesh, int32_t* ibc,
                     int32_t* idonor, double* frac);
int32_t ot2d_ihc_free(ot2d_ihc_lists* l);

/* Installs connectivity for time instance nsp (0-based) of every mesh of the
 * case: the iblank fields and the interpolation lists of later steps. */
int32_t ot2d_set_connectivity(ot2d_case* c, int32_t nsp, const ot2d_ihc_lists* l);
/* Gathers every instance's grid from the device, runs ot2d_ihc_connect with
 * ihcbc[nmesh] and installs the result (src/Turns_overset.f:294-353).
 * *nfringe_total (may be NULL) receives the fringe cells of all meshes.  The
 * ihcbc rows are kept: a pitching case cuts again with them every step. */
int32_t ot2d_connect(ot2d_case* c, const ot2d_ihcbc* ihcbc, int32_t* nfringe_total);
/* do_interpolations once, outside a step. */
int32_t ot2d_interpolate(ot2d_case* c);

/* Error text of the last failing call on this thread. */
const char* ot2d_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* OT2D_H */
