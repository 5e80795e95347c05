/*
 * libfribo_gpu -- MI355X (gfx950) fractional-step fluid step of FRiBoIS2D.
 *
 * C ABI (plain pointers and sizes) that replaces the body of the reference's
 * per-time-step routine
 *
 *     SUBROUTINE fluid_tsch(ns,str)            src/fluid_tsch.F90:4-204
 *
 * together with everything it calls on the fluid side (cfl, boucdq/boucqt,
 * hdnl2/hdnl3, invtr2/invtr3, lgrdistribution, forcing_MLS, helsolve and the
 * solq* sweeps, divg, phcalc_outflow, updvp, prcalc, pr_submean).  The
 * Fortran driver keeps its module state in host arrays with the reference
 * extents; it uploads them once, calls fribo_fluid_tsch once per step and
 * copies back only what its host consumers read.  The Fortran interface
 * module is fribois2d_amd/fortran/fribo_gpu_mod.f90.
 *
 * Host array layouts are exactly the reference's (Fortran column-major, one
 * MPI rank so kstart=1, kend=n3m):
 *   q2, q3, pr          (1:n2, -2:n3m+3)     mpi_routines.F90:428-441
 *   ru2, ru3            (1:n2, 1:n3m)        mpi_routines.F90:452-458
 *   qb2s ... dq3x2o     (1:n2)               utils_proc.F90:277-284
 *   lmcc, lmvel         (nlm, 2)  [y, z]     declarations.F90:258-262
 *   lmA                 (nlm)
 *   lmind               (nlm, 2, 2)          lgrdistribution.F90:46-72
 *   eulins              (nlm, 2, 2, 9)       forcing_MLS.F90:94
 *   phi, phipred        (nlm, 2, 9)          forcing_MLS.F90:69,110
 *   cl                  (nlm, 2)             forcing_MLS.F90:81
 *   fey, fez            (1:n2, -2:n3m+2)     utils_proc.F90:310-320
 *   zmm, zmmt           (ldz, *) leading dimension ldz (n3 in the reference)
 *
 * Every call returns 0 on success or one of the FRIBO_ERR_* codes; the
 * codes that the reference turns into stopcase/MPI_ABORT are listed with the
 * reference line that raises them.  fribo_last_error() gives a message.
 * All calls of one handle must come from one host thread.
 */
#ifndef FRIBO_GPU_H
#define FRIBO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FRIBO_OK = 0,
    FRIBO_ERR_DT_TOO_SMALL = 1,     /* fluid_tsch.F90:43-47   dt < 1e-8            */
    FRIBO_ERR_CFL_TOO_LARGE = 2,    /* fluid_tsch.F90:50-54   cflm*dt > cfllim     */
    FRIBO_ERR_MARKER_OUTSIDE = 3,   /* lgrdistribution.F90:32-36                   */
    FRIBO_ERR_PARTITION_UNITY = 4,  /* MLSint.F90:157-163     |1-sum(phi)| > 1e-10 */
    FRIBO_ERR_ARG = 5,              /* invalid argument / handle                   */
    FRIBO_ERR_HIP = 6,              /* HIP runtime failure                         */
    FRIBO_ERR_LIMIT = 7             /* size beyond what this build supports        */
};

/* Scalars of fluid_par.in / global_par.in that the fluid step reads
 * (utils_proc.F90:59-83,142-170).  Logicals are 0/1. */
typedef struct {
    int32_t n2, n3;
    int32_t nsst, idtv, ibact, forst, wfun, inslws, inslwn;
    double alx3, rext2, visc, cou, q3uw, q3lw, dtmax, cflmax, cfllim, radiusSD;
} fribo_params;

typedef struct fribo_ctx* fribo_handle;

/* Set-up: fluid_pre.F90:27-70 scalars, indic/cordin/coetar/tridiag_matrices
 * metrics (indic.F90:4-44, cordin.F90:105-215, coetar.F90:5-144,
 * phini.F90:109-145) from the node coordinates rc(1:n2) and zz(1:n3), and
 * the phini eigen-decomposition zmm, zmmt, wr (phini.F90:35-92) computed by
 * the host; pass zmm = zmmt = wr = NULL to have fribo_phini's decomposition
 * computed here instead.  Allocates all device memory. */
int fribo_init(const fribo_params* p, const double* rc, const double* zz, const double* zmm,
               const double* zmmt, int ldz, const double* wr, fribo_handle* out);
int fribo_finalize(fribo_handle h);

/* Copy the prognostic state host -> device / device -> host (reference
 * layouts above).  NULL pointers are skipped. */
int fribo_set_state(fribo_handle h, const double* q2, const double* q3, const double* pr,
                    const double* ru2, const double* ru3, const double* qb2s, const double* qb3s,
                    const double* qb2n, const double* qb3n, const double* dq2x2o,
                    const double* dq3x2o);
int fribo_get_state(fribo_handle h, double* q2, double* q3, double* pr, double* ru2, double* ru3,
                    double* qb2s, double* qb3s, double* qb2n, double* qb3n, double* dq2x2o,
                    double* dq3x2o);

/* Marker positions / velocities / lengths as rigid() leaves them
 * (rigid.F90:4-138, read_rigid utils_proc.F90:560-565). */
int fribo_set_markers(fribo_handle h, int32_t nlm, const double* lmcc, const double* lmvel,
                      const double* lmA);

/* One time step: fluid_tsch.F90:25-166.  dt is in/out (CFL control may
 * change it), dt_o is the previous step's dt (used by nsst=1), cflm out. */
int fribo_fluid_tsch(fribo_handle h, int32_t ns, double* dt, double dt_o, double* cflm);

/* The same step in two parts, for stage-by-stage checks: CFL and dt control
 * (fluid_tsch.F90:30-57), then one substep n=1..nsst (fluid_tsch.F90:62-166). */
int fribo_cfl_dt(fribo_handle h, int32_t ns, double* dt, double dt_o, double* cflm);
int fribo_substep(fribo_handle h, int32_t n, int32_t ns);
/* fribo_substep in the launch sequence fribo_fluid_tsch runs: substep n's
 * boundary kernel also applies substep n-1's pressure mean and boundary
 * update (pr_submean + boucqt, updvp.F90:112-144, iobcs.F90:81-108), and only
 * the last substep ends with the mean kernel.  Between two substeps "pr"
 * still holds the pending mean (a constant) and qb2s..qb3n their pending
 * increments dqb*; the other arrays are final. */
int fribo_substep_folded(fribo_handle h, int32_t n, int32_t ns);

/* Named arrays in the reference layout: "q2","q3","pr","dph","ru2","ru3",
 * "rhs2","rhs3" (rhs2/rhs3 include the forcing term f*dt at the IB
 * support cells, i.e. the rhs helsolve.F90:22,47 builds), "fey","fez",
 * "qb2s","qb3s","qb2n","qb3n","dqb2s","dqb3s","dqb2n","dqb3n","dq2x2o",
 * "dq3x2o", the 1-D metric arrays ("rc","rm","zz","zm","g2rc","g2rm",
 * "g3rc","g3rm","udx2c","udx2m","udx3c","udx3m","am2j","ac2j","ap2j",
 * "am3j","ac3j","ap3j","am3ck","ac3ck","ap3ck","am3sk","ac3sk","ap3sk",
 * "amphj","acphj","apphj","amphk","acphk","apphk"), and the marker arrays
 * "lmind","eulins" (int32), "phi","phipred","cl".  "scalars" returns
 * {dt, beta, al, ga, ro, cflm, q3avg, qinf, qout, dx2, dx3, ren}. */
long fribo_array_size(fribo_handle h, const char* name);
int fribo_get_array(fribo_handle h, const char* name, void* out, long n);

int fribo_sync(fribo_handle h);

/* Windowed read-back for host consumers that need only some planes (output
 * slices, restart in pieces, probes): planes kfirst..klast (reference k
 * indices, -2..n3m+3) of the fields selected by mask (1 q2, 2 q3, 4 pr,
 * 8 ru2, 16 ru3) into the host arrays of reference extents; ru2/ru3 take the
 * window clipped to 1..n3m; the other planes of the host arrays are left as
 * they are.  Same values as fribo_get_state on those planes. */
int fribo_fetch_fields(fribo_handle h, int32_t mask, int32_t kfirst, int32_t klast, double* q2, double* q3,
                       double* pr, double* ru2, double* ru3);

/* Loads on the body after a step: SUBROUTINE loads_transfer_c(ns)
 * (loads_transfer.F90:4-162, called at fluid_tsch.F90:175-177).  Probes at
 * ddmin*kprobe^2 along the markers' normals (ddmin = smallest node spacing,
 * cordin.F90:217-233; kprobe from global_par.in), MLS shape functions with
 * derivatives (MLS_SF_d, MLSint.F90:173-258), pressure press(nlm) and
 * friction stress tau(nlm,2) at every marker -- the press_r / tau_r that
 * rigid() integrates.  lmnor, lmacc (nlm,2) as rigid() leaves them
 * (rigid.F90:104,123).  Uses the markers of the last step.  Returns
 * FRIBO_ERR_PARTITION_UNITY when a probe's MLS fit fails
 * (MLSint.F90:249-255). */
int fribo_loads_transfer(fribo_handle h, double kprobe, const double* lmnor, const double* lmacc,
                         double* press, double* tau);

/* No counterpart in the reference: fribo_loads_transfer for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * coupling loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results; every device buffer of the call lives in a static device
 * array of the library (1.8 MB per device, held until the process ends), so
 * it serves at most 8192 markers and returns FRIBO_ERR_LIMIT above that.
 * Calls on one device share that array: calls that overlap on one GPU (two
 * host threads, two handles on one device) race and must be serialised by
 * the caller. */
int fribo_loads_transfer_static(fribo_handle h, double kprobe, const double* lmnor, const double* lmacc,
                                double* press, double* tau);

/* The per-step diagnostics of SUBROUTINE ttime (ttime.F90:29-69,136) on the
 * device, so that the fields need not cross PCIe for them: out[9] =
 * {vmaxv(1), vmaxv(2) (vcheck.F90:3-36), vminv(1), vminv(2)
 * (vcheck.F90:42-73), divgck (divergence.F90:49-78), kinen
 * (utils_proc.F90:474-515), getintvort (utils_proc.F90:582-629), cflm and
 * q3avg of cfl (cfl.F90:4-51)}.  Maxima and minima are exact; the sums are
 * fixed-order tree sums. */
int fribo_diagnostics(fribo_handle h, double* out);

/* Rollback copy for iterated (strong) fluid-structure coupling, the role of
 * the q2_o, q3_o, pr_o, ru2_o, ru3_o arrays the reference declares "for
 * strong coupling" but never uses (declarations.F90:155-159):
 * fribo_save_state copies the prognostic state (fields, outflow arrays, step
 * scalars) on the device;
 * fribo_restore_state puts it back, so the next fribo_fluid_tsch repeats the
 * saved step with new marker kinematics.  One slot per handle. */
int fribo_save_state(fribo_handle h);
int fribo_restore_state(fribo_handle h);

/* SUBROUTINE checkbcs (ckbcs.F90:5-86, called at fluid_tsch.F90:169): the
 * marker velocity error after the step, bcerrv/bcerrw (nlm) (NULL to skip),
 * and nbcv/nbcw = norm2(err)/nlm. */
int fribo_checkbcs(fribo_handle h, double* bcerrv, double* bcerrw, double* nbcv, double* nbcw);

/* Execution option (no reference counterpart): with on != 0 the step's
 * kernels are captured once into a HIP graph and replayed by every
 * fribo_fluid_tsch call (re-captured when the marker count changes);
 * results are identical to the default stream launches. */
int fribo_set_graph(fribo_handle h, int32_t on);
const char* fribo_last_error(void);

/* Diagnostic entry point (no reference counterpart): C = A * B^T in FP64 on
 * device buffers of the library (used by the GEMM parity test).  A is MxK
 * (lda), B is NxK (ldb), C is MxN (ldc), all column-major host arrays. */
int fribo_test_gemm(int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                    int32_t ldb, double* C, int32_t ldc);

/* Diagnostic entry point (no reference counterpart): C = A * B in FP64, the
 * form of the transforms along y.  A is MxK (lda), B is KxN (ldb >= K), C is
 * MxN (ldc), all column-major host arrays. */
int fribo_test_gemm_nn(int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                       int32_t ldb, double* C, int32_t ldc);

/* Execution option (no reference counterpart): the axis of the Poisson
 * eigen-transform.  axis 0: along z with the caller's zmm/zmmt/wr, the
 * reference's order (phcalc_outflow.F90:4-71); axis 1: along y with the
 * eigen-system of the y pressure operator built by fribo_init (one
 * tridiagonal system along z per y-mode), available when n2-1 < n3-1 and
 * both 1-D operators have their Neumann null mode; axis -1: y when
 * available, else z (the default).  FRIBO_ERR_ARG when it would change the
 * axis of an instance that is in a batch.  Both orders solve the same discrete
 * system with the same gauge (the constant fixed by the z-order's pinned
 * null mode); the y order costs 2*(n3-1)*(n2-1)^2 flops per transform
 * instead of 2*(n2-1)*(n3-1)^2.  fribo_get_transform returns the axis in
 * use. */
int fribo_set_transform(fribo_handle h, int32_t axis);
int fribo_get_transform(fribo_handle h);

/* Execution option (no reference counterpart): the mirror form of the
 * transform along y.  When the y grid is symmetric about the channel
 * centre (the y pressure operator equals its mirror image to 1e-10 of its
 * largest coefficient) and n2-1 is even, every eigenvector of the operator
 * is even or odd about the centre; with the pairs (y_j, y_{n-1-j}) replaced
 * by their sum and difference before the forward transform and after the
 * backward one, both transforms are block-diagonal with two halves, half
 * the flops of the dense form.  The two halves are the eigen-systems of the
 * top half's operator with the centre coupling added or subtracted.
 * on = 1 selects it (the default where it is available), 0 the dense form;
 * the results agree to rounding.  It applies to the FP64 transforms along y
 * (single-instance and batched steps; the FP32 transforms of
 * fribo_set_precision keep the dense form).  FRIBO_ERR_LIMIT for on = 1
 * without a symmetric grid, FRIBO_ERR_ARG for an instance that is in a
 * batch.  fribo_get_ymirror returns 1 when the mirror form is selected. */
int fribo_set_ymirror(fribo_handle h, int32_t on);
int fribo_get_ymirror(fribo_handle h);

/* Execution option (no reference counterpart): the form of the k-direction
 * line solves (solq2k / solq3k, solve_fluid.F90:107-255).  form 1: one
 * pass per field that reads the right-hand side once (chunks of at most 32
 * rows, i.e. k-lines of at most 2048 rows; the default where it fits);
 * form 0: three passes (chunk elimination, separator system, chunk
 * back-substitution), any length.  Both give identical results.
 * FRIBO_ERR_LIMIT for form 1 on longer lines, FRIBO_ERR_ARG for an instance
 * that is in a batch.  fribo_get_ksolve returns the form in use. */
int fribo_set_ksolve(fribo_handle h, int32_t form);
int fribo_get_ksolve(fribo_handle h);

/* Execution option (no reference counterpart; the reference's DGEMMs are
 * FP64): precision of the Poisson eigen-transforms along y.  refine < 0:
 * FP64 MFMA (the default).  refine = n in 0..3: the transforms on FP32 MFMA
 * with FP32 accumulation (operands rounded to FP32, results back in FP64),
 * the per-mode line solves in FP64, then n sweeps of FP64-residual iterative
 * refinement x += solve(rhs - (A_y + A_z) x).  FRIBO_ERR_LIMIT when the
 * transform is along z.  fribo_get_precision returns the setting. */
int fribo_set_precision(fribo_handle h, int32_t refine);
int fribo_get_precision(fribo_handle h);
/* Diagnostic entry point: the FP32 transform C = A * B on host arrays. */
int fribo_test_gemm32(int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                      int32_t ldb, double* C, int32_t ldc);

/* phini (phini.F90:4-99) without LAPACK: the z pressure operator of
 * tridiag_matrices (phini.F90:121-142) built from zz(1:n3) and alx3 is
 * similar to a symmetric tridiagonal matrix, decomposed by implicit QL;
 * zmm (eigenvectors as columns), zmmt = zmm^-1 (n3-1 x n3-1, leading
 * dimension ld) and wr(n3-1).  Same eigen-pairs as DGEEV up to order and
 * scaling, which the Poisson solve does not depend on; O(n^2) + a threaded
 * O(n^3) eigenvector accumulation instead of DGEEV's dense O(n^3). */
int fribo_phini(int32_t n3, const double* zz, double alx3, int32_t ld, double* zmm, double* zmmt, double* wr);

/* fribo_phini on the GPU (phini.F90:4-99; no reference counterpart runs it on
 * a device): the same operator and similarity, the symmetric tridiagonal
 * eigen-problem solved by rocSOLVER's divide-and-conquer STEDC (eigenvalues
 * in increasing order), Y and Y^-1 assembled by a kernel; host arrays in
 * and out.  For many instances of uneven grids, one call per instance. */
int fribo_phini_device(int32_t n3, const double* zz, double alx3, int32_t ld, double* zmm, double* zmmt,
                       double* wr);

/* Diagnostic entry point (no reference counterpart): eigen-decomposition of
 * an n x n tridiagonal operator with rows (am, ac, ap), 0-based, that is
 * similar to a symmetric one (am[j+1]*ap[j] > 0): Y (eigenvectors as
 * columns), Yinv = Y^-1 (both n x n, column-major) and the eigenvalues wy. */
int fribo_test_ytransform(int32_t n, const double* am, const double* ac, const double* ap, double* Y, double* Yinv,
                          double* wy);

/* Diagnostic entry point (no reference counterpart): the mirror form of the
 * same decomposition for an operator symmetric about its centre (n even):
 * Ys (the backward transform on the pair sums and differences), Ysi (the
 * forward one), both block-diagonal with two n/2 x n/2 blocks, and the
 * eigenvalues wys in their mode order.  Returns 7 when the operator is not
 * symmetric to within tol times its largest diagonal entry. */
int fribo_test_ymirror(int32_t n, const double* am, const double* ac, const double* ap, double tol, double* Ys,
                       double* Ysi, double* wys);

/* Diagnostic entry point (no reference counterpart): `reps` streaming copies
 * of `bytes` (a multiple of 8) between two device buffers after one warm-up
 * copy; returns the mean device time per copy.  It is the known-size kernel
 * that scales the HBM traffic counters of the step's kernels. */
int fribo_test_copy(int64_t bytes, int32_t reps, double* ms_per_copy);

/* Diagnostic timing (no reference counterpart): with profiling on (on = 1),
 * every kernel launch of the step is bracketed by HIP events on the
 * library's stream; on = 2 brackets only the transform GEMMs (the least
 * intrusive timing of the dominant kernel).  fribo_get_profile returns, per
 * kernel class, the summed device time (ms) and the number of launches since
 * the last reset. */
#define FRIBO_NPROF 12
int fribo_set_profiling(fribo_handle h, int32_t on);
int fribo_get_profile(fribo_handle h, double* ms, int32_t* launches, int32_t nclass);
int fribo_reset_profile(fribo_handle h);
const char* fribo_profile_name(int32_t cls);

/* Batches of independent instances (no counterpart in the reference: many
 * runs of fluid_tsch -- sweeps over Re, reduced velocity, mass ratio -- on
 * uneven grids, advanced together).  The instances are ordinary handles
 * made by fribo_init and filled by the calls above; fribo_batch_create
 * groups them: from then on they share one stream, and fribo_batch_step
 * advances all of them by one time step (fluid_tsch.F90:25-166 each) with
 * one kernel launch per stage for the whole batch.  Each instance computes
 * exactly what its own fribo_fluid_tsch would, up to the rounding of the
 * Poisson gauge sums when the batch's transform tiles differ from the
 * instance's own (1e-15 relative).  Every instance takes the transform
 * along y when all of them can, else along z.
 *   fribo_batch_set_markers: the marker kinematics of every instance in one
 *     copy, lmcc_i, lmvel_i, lmA_i of instance i at offset 5*sum(nlm_<i)
 *     (the counts of each instance's last fribo_set_markers);
 *   fribo_batch_step: dt[i] in/out, dt_o[i] (NULL: dt), cflm[i] and err[i]
 *     out (NULL to skip): the codes fribo_fluid_tsch would return;
 *   fribo_batch_set_graph: replay the batch's launches as one HIP graph;
 *   fribo_batch_set_gemm_shape: tile shape of the transforms (0: 32x32,
 *     1: 64x64, 2: 96x128, 3: direct-to-LDS 96x128, 4: direct-to-LDS
 *     96x64 with 16-deep K-slices, -1: default), and
 *     fribo_batch_get_gemm_shape the one in use.
 * Destroy the batch before finalizing its instances. */
typedef struct fribo_batch_ctx* fribo_batch;
int fribo_batch_create(const fribo_handle* h, int32_t n, fribo_batch* out);
int fribo_batch_destroy(fribo_batch b);
int fribo_batch_set_markers(fribo_batch b, const double* mk);
int fribo_batch_step(fribo_batch b, int32_t ns, double* dt, const double* dt_o, double* cflm, int32_t* err);
int fribo_batch_set_graph(fribo_batch b, int32_t on);
int fribo_batch_set_gemm_shape(fribo_batch b, int32_t cfg);
int fribo_batch_get_gemm_shape(fribo_batch b);
/* Autotuner (no reference counterpart): times the batch's forward transform
 * with every tile shape all its instances can run (reps launches each, no
 * change to any state) and keeps the fastest, counting with the transform
 * along y the chunk pass of the per-mode z-lines where the shape leaves it
 * to its own kernel instead of the transform's epilogue; ms[10] gets the
 * time per launch of shapes 0..4 without (ms[0..4]) and with (ms[5..9])
 * that pass (0: not available), best the shape kept. */
int fribo_batch_autotune(fribo_batch b, int32_t reps, double* ms, int32_t* best);

/* Learned first iterate of the strong coupling loop (no counterpart in the
 * reference, whose coupling is loose; its unused iterated-coupling state is
 * declarations.F90:250,271-278): an MLP ni -> 64 tanh -> 64 tanh -> no with
 * FP32 parameters [W1 (64 x ni, row-major), b1, W2 (64 x 64), b2,
 * W3 (no x 64), b3] (fribo_mlp_size floats), trained by the library's own
 * forward, backward and Adam kernels (FP32, gradients summed over a
 * minibatch in sample order) and run in BF16 with FP32 accumulation.
 * Samples are row-major (n x ni inputs, n x no targets).  Codes are the
 * FRIBO_ERR_* values. */
typedef struct fribo_mlp_ctx* fribo_mlp;
int fribo_mlp_create(int32_t ni, int32_t no, const float* init, uint64_t seed, fribo_mlp* out);
int fribo_mlp_destroy(fribo_mlp m);
int fribo_mlp_size(fribo_mlp m);
int fribo_mlp_get_params(fribo_mlp m, float* out);
int fribo_mlp_grad(fribo_mlp m, int32_t n, const float* X, const float* T, float* grad, float* loss);
int fribo_mlp_train(fribo_mlp m, int32_t n, const float* X, const float* T, int32_t epochs, int32_t batch, float lr,
                    float* loss);
int fribo_mlp_predict(fribo_mlp m, int32_t n, const float* X, float* Y);

#ifdef __cplusplus
}
#endif
#endif
