// Internal definitions shared by the gfx950 kernels of libfribo_gpu.
//
// Device layout: every 2-D field uses the reference's logical extents
// (1:n2, -2:n3m+3) with j fastest and a padded row pitch P (multiple of 16
// doubles, 128-byte rows), so element (j,k) lives at (j-1) + (k+2)*P.
// 1-D grid arrays are 1-based (index 0 unused) exactly like the Fortran
// module arrays they mirror.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FRIBO_DIMSD 9
#define FRIBO_KMAXCH 64      // max chunks of a partitioned k-line (one wave)
// CFL reduction: one partial (max, q3 sum) per workgroup of k_cfl, which
// covers 256 columns and FRIBO_CFL_KT planes
#define FRIBO_CFL_KT 8
inline int fribo_ncfl(int n2m, int n3m) { return ((n2m + 255) / 256) * ((n3m + FRIBO_CFL_KT - 1) / FRIBO_CFL_KT); }
#define FRIBO_LWORK 20       // per-marker record of the load computation
#define FRIBO_STATIC_LOADS_MAX_NLM 8192   // markers fribo_loads_transfer_static serves

// Marching tiles of the stencil kernels (k_rhs, k_div, k_upd): a wave owns
// FRIBO_WJ output columns (plus a halo lane each side) and MK planes
// (FRIBO_MK_RHS for k_rhs, whose register window limits its occupancy;
// FRIBO_MK for k_div and k_upd, whose shorter tiles put more waves in
// flight).  k_upd writes one pressure partial per tile.
#ifndef FRIBO_MK
#define FRIBO_MK 8
#endif
#ifndef FRIBO_MK_RHS
#define FRIBO_MK_RHS 16
#endif
#define FRIBO_WJ 62
__host__ __device__ inline int fribo_npr_part(int n2m, int n3m)
{
    return ((n2m + FRIBO_WJ - 1) / FRIBO_WJ) * ((n3m + FRIBO_MK - 1) / FRIBO_MK);
}

struct DevScalars {
    // step inputs, written by the host before every step (one 24-byte copy)
    double in_dt, in_dto;
    int in_ns;
    int err;                 // first error code raised on the device (host zeroes it)
    double dt, dt_o, beta, al, ga, ro, alre, betadx;
    double cflm, q3avg;
    double pend;             // pressure mean not yet subtracted from the interior
    double goff;             // sum of the means not yet subtracted from the pressure
                             // entries no kernel reads (row n2, ghost planes)
    double dphoff;           // constant to subtract from the interior of dph (gauge
                             // of the transform along y; 0 along z)
    double qinf, qout, area;
    double ugkks, ugkkn;
};

// Precomputed, line-invariant factors of one partitioned tridiagonal system
// (rows r = 0..n-1, normalised diagonal 1).  Chunk p covers rows
// [cs[p], cs[p+1]); its last row is the separator X_p, the others are its
// interior rows.  Row arrays are stored either by row (k-direction kernels,
// whose waves read one row at a time) or lane-interleaved, row t of chunk p
// at [t*64 + p] (j-direction kernels, whose lanes own the chunks).
struct KSys {
    int n;                   // rows of the system
    int nch;                 // chunks (<= FRIBO_KMAXCH)
    int kofs;                // k (or j) index of row 0
    int maxm;                // largest chunk length
    int inter;               // 1: lane-interleaved row arrays
    int src;                 // coefficient source: 0 am3sk.., 1 am2j.., 2 am3j..
    int* cs;                 // [nch+1] chunk starts
    double* a;               // sub-diagonal (0 at row 0)
    double* c;               // super-diagonal (0 at row n-1)
    double* S;               // rhs scaling 1/(1 - betadx*ac)
    double* cp;              // forward multipliers c'
    double* inv;             // 1/pivot
    double* vp;              // forward left-coupling vector v'
    double* g;               // weights giving y_s from y' (sum g_i y'_i)
    double* wl;              // [64] w' at the last interior row of each chunk
    double* k1;              // [6*64] PCR multipliers (lower)
    double* k2;              // [6*64] PCR multipliers (upper)
    double* ib;              // [64] 1/beta after PCR
    double* bc;              // [2] couplings of rows 0 / n-1 to the Dirichlet ends
};

// Fixed partition of the Poisson j-lines (n2m rows; the per-mode matrices
// are factored in registers by the line kernel).  Coefficients of row t of
// chunk p are lane-interleaved at [t*64 + p].
struct PSys {
    int n, nch, maxm, stride;
    int* cs;                 // [nch+1]
    double *am, *ac, *ap;    // amphj, acphj, apphj, interleaved
};

struct FluidDev {
    int n2, n3, n2m, n3m;
    int P;                   // row pitch (doubles)
    int nk;                  // planes (n3m+6)
    int nsst, idtv, ibact, wfun, inslws, inslwn;
    int nlm;
    int null_mode;           // 1-based index of the singular Poisson mode, 0 if none
    int taxis;               // eigen-transform axis of the Poisson solve: 0 z (zmm), 1 y (ymm)
    int refine;              // -1: FP64 transforms; n >= 0: FP32 transforms + n FP64-residual sweeps
    double alx3, rext2, ren, visc, cou, q3uw, q3lw, dtmax, cflmax, cfllim, radiusSD;
    double dx2, dx3, dx2q, dx3q;
    double gam[4], rom[4], alm[4];
    // fields (ghosted layout)
    double *q2, *q3, *pr, *dph, *ru2, *ru3, *rhs2, *rhs3, *sA, *sB;
    // Poisson transform
    double *hmn;             // n2m x n3m, leading dimension P
    double *zmm, *zmmt;      // n3m x n3m, leading dimension ldz
    int ldz;
    // transform along y: Y (eigenvectors of A_y), Yinv, eigenvalues wy (0-based),
    // the y null mode (1-based), gauge weights zg = zmmt(null_mode, :) and 1/sum(zg),
    // and the per-mode z-line workspace (7 x n2m x 64)
    double *ymm, *ymmi, *wy, *zg, *ysum;
    int ldy, ynull;
    double zginv;
    // mirror form of the transform along y (grids symmetric about the channel
    // centre): block-diagonal Ys (backward), Ysi (forward), the eigenvalues of
    // its mode order and its null mode; ysh = n2m / 2 when in use, 0 otherwise
    double *ysm, *ysmi, *wys;
    int ysh, ynulls;
    double* gpart;           // gauge partials of the backward transform along y, one per tile column
    int ngpart;
    // outflow arrays (1-based, n2+2)
    double *qb2s, *qb3s, *qb2n, *qb3n, *dqb2s, *dqb3s, *dqb2n, *dqb3n, *dq2x2o, *dq3x2o;
    // 1-D grid (1-based)
    double *rc, *rm, *zz, *zm, *g2rc, *g2rm, *g3rc, *g3rm, *udx2c, *udx2m, *udx3c, *udx3m;
    double *am2j, *ac2j, *ap2j, *am3j, *ac3j, *ap3j;
    double *am3ck, *ac3ck, *ap3ck, *am3sk, *ac3sk, *ap3sk;
    double *amphj, *acphj, *apphj, *amphk, *acphk, *apphk, *wr;
    int *jmv, *jpv, *kmv, *kpv;
    // reductions
    double *cfl_part;        // [2*ncfl]
    int ncfl;                // CFL partials (workgroups of k_cfl)
    double *pr_part;         // [npr_part]
    int npr_part;            // partials of k_upd (those of the mirror form's k_upd: npr_count)
    // Helmholtz systems: ks[0] q2 along k, ks[1] q3 along k,
    // js[0] q2 along j, js[1] q3 along j; Poisson j-lines
    KSys ks[2];
    KSys js[2];
    PSys ps;
    int jstride;             // LDS chunk stride of the j-line kernels (odd)
    int kform;               // k-line solve: 1 one fused pass (chunks of <= 32 rows), 0 three passes
    int lds_max;             // LDS bytes a workgroup may use on this device
    double *kYs, *kYe, *kDe, *kX;  // [2][n2m][64] reduced-system data (chunk fastest)
    // markers
    double *lmcc, *lmvel, *lmA;    // (nlm,2) column-major
    int *lmind;                    // (nlm,2,2)
    int *eulins;                   // (nlm,2,2,9)
    double *phi, *phipred, *cl;    // (nlm,2,9), (nlm,2)
    int *active;                   // [nlm]
    double *contrib;               // [2][nlm*9]
    unsigned long long *keys;      // [2][maxkeys]
    double *segval;                // [2][maxkeys] cell value at segment heads
    int maxkeys;
    // loads on the body (loads_transfer_c) and the forcing check (checkbcs)
    double *lmnor, *lmacc;         // (nlm,2) column-major, from the host
    double *press, *tau;           // (nlm), (nlm,2)
    double *lwork;                 // [nlm][FRIBO_LWORK] per-marker probe data
    double *bcerr;                 // (nlm,2)
    DevScalars* S;
};

// s_barrier without the memory fence of __syncthreads (whose release waits
// for every outstanding global access of the wave, loads into LDS and
// stores included); the asm memory clobber keeps the compiler from moving
// memory accesses across it.
__device__ inline void raw_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Workgroup barrier that orders LDS only: waits for this wave's LDS
// accesses (lgkmcnt(0)), not for its global stores, which __syncthreads
// would drain first.  For exchanges through LDS whose global stores no
// other thread of the workgroup reads back.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
    raw_barrier();
}

// Value of v in lane l of the wave (l wave-uniform): a matrix row that one
// lane loaded and every lane of the wave uses, moved through scalar registers
// instead of a scalar memory load per use.
__device__ inline double lane_bcast(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 1/x from the hardware estimate refined by two Newton steps (within an ulp
// of the rounded quotient for the pivots of the diagonally dominant line
// systems, far from zero and from the denormal range).  A fraction of the
// latency of the IEEE division sequence, which matters on the serial
// elimination chains.
__device__ inline double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__host__ __device__ inline size_t fidx(const FluidDev& F, int j, int k)
{
    return (size_t)(j - 1) + (size_t)(k + 2) * (size_t)F.P;
}

// Error codes (match include/fribo_gpu.h)
enum {
    FRIBO_E_OK = 0,
    FRIBO_E_DT_TOO_SMALL = 1,
    FRIBO_E_CFL_TOO_LARGE = 2,
    FRIBO_E_MARKER_OUTSIDE = 3,
    FRIBO_E_PARTITION_UNITY = 4,
    FRIBO_E_ARG = 5,
    FRIBO_E_HIP = 6,
    FRIBO_E_LIMIT = 7,
};

// Kernel launchers (each .hip file owns its kernels).
void launch_cfl(const FluidDev& F, hipStream_t st);
void launch_bc(const FluidDev& F, int n, bool mean, hipStream_t st);
void launch_rhs(const FluidDev& F, hipStream_t st);
void launch_div(const FluidDev& F, hipStream_t st);
void launch_upd(const FluidDev& F, hipStream_t st);
void launch_mean(const FluidDev& F, hipStream_t st);
void launch_mls_setup(const FluidDev& F, hipStream_t st);
void launch_mls_force(const FluidDev& F, hipStream_t st);
void launch_jsolve_helm(const FluidDev& F, hipStream_t st);
void launch_jsolve_poisson(const FluidDev& F, hipStream_t st);
void launch_ksolve(const FluidDev& F, hipStream_t st);
// true when both k-systems have chunks of at most 32 rows (the one-pass form)
bool ksolve_fused_fits(const FluidDev& F);
// First pass of the y-order z-lines (chunk elimination, fribo_ychunk.h) run
// by the epilogue of the forward transform on its own output tile, when
// every 64-column tile holds whole chunks of the line partition (csz rows
// each, csz - 1 <= 32).  csz = 0: no fused pass.
struct YFuse {
    const double *wy, *amphk, *acphk, *apphk;
    double* ysum;
    int n2m, csz;
};

// e0/e1 (optional): events the dispatch records at its start and end
void launch_gemm_nt(int M, int N, int K, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc, const DevScalars* S, hipStream_t st, hipEvent_t e0 = nullptr,
                    hipEvent_t e1 = nullptr);
// Returns true when the fused chunk pass of yf ran (yf->csz > 0 and the
// direct-to-LDS kernel with 64-column tiles took the launch).
bool launch_gemm_nn(int M, int N, int K, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc, const DevScalars* S, hipStream_t st, const double* zg = nullptr,
                    double* gpart = nullptr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                    const YFuse* yf = nullptr, int ks = 0);
int gemm_nn_tiles_n(int M, int N, int K, const double* A, int lda, const double* B, int ldb);
// first_done: the chunk pass already ran in the forward transform's epilogue
void launch_ysolve(const FluidDev& F, hipStream_t st, bool first_done = false);
void launch_loads(const FluidDev& F, double hpr, hipStream_t st);
// mixed-precision Poisson solve: r = rhs - (A_y + A_z) x (r may alias rhs),
// and the FP32 transform C = X0 + A * B (X0 optional, may alias C)
void launch_pres_residual(const FluidDev& F, const double* x, const double* rhs, double* r, hipStream_t st);
void launch_gemm32(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                   const double* X0, const DevScalars* S, hipStream_t st, const double* zg, double* gpart,
                   hipEvent_t e0, hipEvent_t e1);
int gemm32_tiles_n(int N);
void launch_checkbcs(const FluidDev& F, double* nb, hipStream_t st);
void launch_diag(const FluidDev& F, double* part, int kv, double* out, hipStream_t st);
// device: the symmetrised eigen-problem on the GPU (rocSOLVER STEDC) instead
// of the host QL
int ytransform_build(int n, const double* am, const double* ac, const double* ap, int ld, double* Y, double* Yinv,
                     double* wy, bool device = false);
// mirror form (block-diagonal halves) when the operator is mirror-symmetric
// to within tol * max|ac|; -1 otherwise
int ytransform_mirror(int n, const double* am, const double* ac, const double* ap, int ld, double* Ys, double* Ysi,
                      double* wys, double tol);
int phini_sym(int n3, const double* zz, double alx3, int ld, double* zmm, double* zmmt, double* wr,
              bool device = false);
int jsolve_stride(const FluidDev& F);
size_t jsolve_lds_bytes(const FluidDev& F);
size_t sysprep_lds_bytes(const FluidDev& F);
