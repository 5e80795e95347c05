// FP64 MFMA GEMM for the Poisson eigen-transforms (phcalc_outflow.F90:31,61).
//
//   C[i + n*ldc] = sum_kk A[i + kk*lda] * B[n + kk*ldb]     (C = A * B^T)
//   C[i + n*ldc] = sum_kk A[i + kk*lda] * B[kk + n*ldb]     (C = A * B, "NN")
//
// Both transforms have this form on the device: the forward transform
// hmn = dph * zmmt^T takes A = dph (j contiguous) and B = zmmt; the backward
// transform dph = V * zmm^T takes A = V and B = zmm in the reference's own
// column-major layout.  Every K-slice of both operands is therefore a set of
// contiguous columns and loads fully coalesced.
//
// v_mfma_f64_16x16x4_f64 fragment layout: A fragment lane l holds
// A[l%16][l/16]; B fragment lane l holds B[l/16][l%16]; accumulator element v
// of lane l is C[l/16 + 4v][l%16].
//
// Tiling (template): BM x BN output tile per workgroup of WM x WN waves, each
// wave owning (BM/WM) x (BN/WN) as 16x16 MFMA blocks; BK = 16 with two LDS
// buffers (one barrier per K-slice: the next slice is fetched into registers
// while the current one is multiplied, then stored into the other buffer).
// LDS rows are padded to 16 (mod 32) doubles so the four k-rows a fragment
// read touches fall in disjoint bank halves.  In the NN form (the transforms
// along y, whose B operand is the field itself, k-contiguous) the B slice is
// read as 16-double runs of each column and stored with a row pitch of BN+1,
// which spreads the transposing LDS stores over the banks.
//
// Large problems (k_gemm_dma): 96 x 128 tiles of 16 waves (48 x 16 each, one
// workgroup per CU; 96 x 64 tiles of 8 waves, two per CU, measured 1-5%
// slower), K-slices of 24 (144 KB of LDS for the three stages; 16-deep
// slices, one more barrier per 48 k, measured 3.5% slower on the 768-deep
// products of the mirror form), and the operand slices move from global
// memory straight into LDS
// (global_load_lds_dwordx4, 16 bytes per lane, no register staging) through
// three stages, so two K-slices of loads are in flight while one is
// multiplied.  Each stage is its own LDS array, the slice loop is unrolled
// by three, and the loads are issued as inline assembly: the compiler then
// neither tracks them nor waits for every outstanding load before a
// fragment read.  Every wave issues the same number of loads per slice
// (addresses clamped into the matrix), so "slice s has landed" is
// vmcnt(loads per slice), waited for explicitly: slice s+1 may stay in
// flight; the workgroup barrier is a plain s_barrier for the same reason.
// The loads of a slice are issued at raised wave priority.
// A direct LDS write lands lane-contiguous, so the LDS layout follows the
// global one: A and the NT-form B as k-rows (pitch 112 / 144 doubles: 48 / 64
// used 16-byte slots of 56 / 72), the NN-form B as n-rows of BK k-values with
// pitch BK+2 (the fragment read then hits 16 distinct bank groups).  Clamped
// loads past the matrix only feed rows and columns that are not stored, or
// the k-rows past K of the last slice, which are cleared before use; when
// the load count of a stage does not divide evenly among the waves, the spare
// loads land past the stage's B block.  The 1536 x 2048 transforms of the
// 2048x1536-cell grid give 256 workgroups, one per CU; workgroups are
// dispatched to the 8 XCDs round-robin, so when the grid is a multiple of 256
// tiles in 8 x 4 patches the tile indices are permuted to give each XCD
// compact patches (its L2 then serves 8 A- and 4 B-tiles to 32 workgroups).
#include "fribo_ychunk.h"
#include "fribo_batch.h"
#include <hip/hip_ext.h>

typedef double f64x4 __attribute__((ext_vector_type(4)));

#define GBK 16
#ifndef GEMM_DMA_BK
#define GEMM_DMA_BK 24
#endif
#ifndef GEMM_DMA_WM
#define GEMM_DMA_WM 2
#endif
#ifndef GEMM_DMA_WN
#define GEMM_DMA_WN 8
#endif
#ifndef GEMM_DMA_BN
#define GEMM_DMA_BN 128
#endif
// tile rows (M) of the 32-tile patch each XCD works on (32 / GEMM_PATCH_M tile
// columns): 8 x 4 patches read 31.3 MiB per mirror-form transform of the
// 2048x1536 grid against 33.8 MiB for 4 x 8 and 53.5 MiB for 2 x 16
#ifndef GEMM_PATCH_M
#define GEMM_PATCH_M 8
#endif

// Gauge partial of the transform along y (zg != nullptr): the tile holding
// row grow of C adds zg(col) * C(grow, col) over its columns, in a fixed
// order (columns of a lane, lanes by butterfly, waves in index order), and
// writes sgn times the sum to *part.  The step then sums the partials in
// tile order.
template <int RM, int RN, int TM, int TN, int NW>
__device__ inline void gauge_partial(const f64x4 (&acc)[RM][RN], int grow, double sgn, int N, int i0, int n0, int wm,
                                     int wn, int lane, int wave, const double* __restrict__ zg, double* part,
                                     double* sh)
{
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < RN; ++c) {
        const int col = n0 + wn * TN + c * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < RM; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = i0 + wm * TM + r * 16 + (lane >> 4) + 4 * q;
                if (row == grow && col < N) v = v + zg[col] * acc[r][c][q];
            }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    if (lane == 0) sh[wave] = v;
    lds_barrier();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int w = 1; w < NW; ++w) t = t + sh[w];
        *part = sgn * t;
    }
}

// Gauge partials of a tile with rows [i0, i0 + BM), tile column by of gy.
// Plain form: C(M-1, :) into part[by].  Mirror form (ks > 0, C before the
// pair butterfly, whose row M-1 is C(0, :) - C(M-1, :)): C(0, :) into
// part[by] and -C(M-1, :) into part[gy + by].
template <int RM, int RN, int TM, int TN, int NW, int BM>
__device__ inline void gauge_partials(const f64x4 (&acc)[RM][RN], int ks, int M, int N, int i0, int n0, int wm,
                                      int wn, int lane, int wave, int by, int gy, const double* __restrict__ zg,
                                      double* part, double* sh)
{
    const bool last = M - 1 >= i0 && M - 1 < i0 + BM;
    if (ks > 0 && i0 == 0) {
        gauge_partial<RM, RN, TM, TN, NW>(acc, 0, 1.0, N, i0, n0, wm, wn, lane, wave, zg, part + by, sh);
        if (last) lds_barrier();
    }
    if (last)
        gauge_partial<RM, RN, TM, TN, NW>(acc, M - 1, ks > 0 ? -1.0 : 1.0, N, i0, n0, wm, wn, lane, wave, zg,
                                          part + (ks > 0 ? gy + by : by), sh);
}

// K range of an output tile with rows [i0, i0 + BM).  Plain form (ks = 0):
// all of K.  Mirror form (ks > 0, A block-diagonal with blocks [0, ks) and
// [ks, K) in both indices): rows below ks take k in [0, ks), rows from ks on
// take [ks, K), a tile holding both takes all of K (the zero blocks add
// exact zeros).  The start is rounded down to a multiple of BK: the k-rows
// it adds meet zeros of A.
__device__ inline void tile_krange(int ks, int K, int i0, int BM, int BK, int& kb, int& ke)
{
    kb = 0;
    ke = K;
    if (ks <= 0) return;
    if (i0 >= ks) kb = ks / BK * BK;
    if (i0 + BM <= ks) ke = ks;
}

template <int BM, int BN, int WM, int WN, bool NN>
__device__ __forceinline__ void gemm_nt_body(int M, int N, int K, const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc, const DevScalars* S, const double* __restrict__ zg, double* gpart, int ks, const VBlock& vb)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int TM = BM / WM, TN = BN / WN;     // wave tile
    constexpr int RM = TM / 16, RN = TN / 16;     // MFMA blocks per wave
    constexpr int LA = BM + ((16 - BM % 32) + 32) % 32;
    constexpr int LB = NN ? BN + 1 : BN + ((16 - BN % 32) + 32) % 32;
    constexpr int NA = (BM * GBK + NT - 1) / NT;  // A elements per thread per slice
    constexpr int NB = (BN * GBK + NT - 1) / NT;
    static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile");
    if (S && S->err) return;
    __shared__ double As[2][GBK][LA];
    __shared__ double Bs[2][GBK][LB];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int i0 = vb.x * BM, n0 = vb.y * BN;
    int kb, ke;
    tile_krange(ks, K, i0, BM, GBK, kb, ke);
    double ra[NA], rb[NB];

    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            int e = t + u * NT;
            int kk = e / BM, r = e % BM;
            bool ok = (e < BM * GBK) && (k0 + kk < ke) && (i0 + r < M);
            ra[u] = ok ? A[(size_t)(k0 + kk) * lda + i0 + r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            int e = t + u * NT;
            int kk = NN ? e % GBK : e / BN, r = NN ? e / GBK : e % BN;
            bool ok = (e < BN * GBK) && (k0 + kk < ke) && (n0 + r < N);
            rb[u] = ok ? (NN ? B[(size_t)(n0 + r) * ldb + k0 + kk] : B[(size_t)(k0 + kk) * ldb + n0 + r]) : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            int e = t + u * NT;
            if (e < BM * GBK) As[buf][e / BM][e % BM] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            int e = t + u * NT;
            if (e < BN * GBK) {
                if (NN) Bs[buf][e % GBK][e / GBK] = rb[u];
                else Bs[buf][e / BN][e % BN] = rb[u];
            }
        }
    };

    f64x4 acc[RM][RN];
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < RN; ++c) acc[r][c] = (f64x4){0.0, 0.0, 0.0, 0.0};

    load(kb);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += GBK) {
        const bool more = k0 + GBK < ke;
        if (more) load(k0 + GBK);
#pragma unroll
        for (int k4 = 0; k4 < GBK; k4 += 4) {
            const int kr = k4 + (lane >> 4);
            double a[RM], b[RN];
#pragma unroll
            for (int r = 0; r < RM; ++r) a[r] = As[buf][kr][wm * TM + r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < RN; ++c) b[c] = Bs[buf][kr][wn * TN + c * 16 + (lane & 15)];
#pragma unroll
            for (int r = 0; r < RM; ++r)
#pragma unroll
                for (int c = 0; c < RN; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < RN; ++c) {
            const int col = n0 + wn * TN + c * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = i0 + wm * TM + r * 16 + (lane >> 4) + 4 * v;
                if (row < M && col < N) C[(size_t)col * ldc + row] = acc[r][c][v];
            }
        }
    if (zg) {
        __shared__ double gsh[WM * WN];
        gauge_partials<RM, RN, TM, TN, WM * WN, BM>(acc, ks, M, N, i0, n0, wm, wn, lane, wave, vb.y, vb.gy, zg,
                                                    gpart, gsh);
    }
}
template <int BM, int BN, int WM, int WN, bool NN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_nt(int M, int N, int K, const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc, const DevScalars* S, const double* __restrict__ zg, double* gpart, int ks)
{
    gemm_nt_body<BM, BN, WM, WN, NN>(M, N, K, A, lda, B, ldb, C, ldc, S, zg, gpart, ks, vb_hw());
}
template <int BM, int BN, int WM, int WN, bool NN>
__global__ __launch_bounds__(64 * WM * WN) void kb_gemm_nt(const BGrid<GemmProb> G)
{
    VBlock vb;
    const GemmProb& P = *bg_block(G, vb);
    gemm_nt_body<BM, BN, WM, WN, NN>(P.M, P.N, P.K, P.A, P.lda, P.B, P.ldb, P.C, P.ldc, P.S, P.zg, P.gpart, P.ks, vb);
}


// Wave priority around the issue of the next slice's loads (2; 1: around
// the MFMA loop instead, 0: none): with two workgroups per CU, a wave that
// has passed the barrier issues its loads ahead of the other workgroup's
// MFMA waves, which keeps the next stages in flight (-2% GEMM time).
#ifndef GEMM_PRIO
#define GEMM_PRIO 2
#endif
#ifndef GEMM_ASM_DMA
#define GEMM_ASM_DMA 1
#endif
// One 16-byte-per-lane global -> LDS load (global_load_lds_dwordx4: lane l
// writes dst + 16 l).  Issued as inline assembly, the compiler does not track
// the LDS write, so it adds no wait for it before later LDS reads of other
// stages; the kernel's own vmcnt wait and barrier order every read after the
// loads of its slice.
__device__ inline void dma16(const double* src, double* dst)
{
#if GEMM_ASM_DMA
    const unsigned lds = __builtin_amdgcn_readfirstlane(
        (unsigned)(uintptr_t)(__attribute__((address_space(3))) double*)dst);
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds), "v"(src) : "memory");
#else
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
#endif
}

template <int BK, bool NN, int WM = 2, int WN = 8, int BN = 128, int BM = 96>
__device__ __forceinline__ void gemm_dma_body(int M, int N, int K, const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc, const DevScalars* S, const double* __restrict__ zg, double* gpart, YFuse yf, int ks, const VBlock& vb)
{
    constexpr int NW = WM * WN;
    constexpr int TM = BM / WM, TN = BN / WN, RM = TM / 16, RN = TN / 16;
    constexpr int LA = BM + 16, SLA = LA / 2;               // A: k-rows, BM / 2 of SLA slots used
    constexpr int LB = NN ? BK + 2 : BN + 16, SLB = LB / 2; // B: n-rows (NN) or k-rows (NT)
    constexpr int SA = BK * LA, SB = NN ? BN * LB : BK * LB;   // doubles per stage
    constexpr int IA = SA / 128, IB0 = SB / 128;            // 1 KB loads per stage
    // every wave issues NI loads per stage; spare loads land past SB
    constexpr int NI = (IA + IB0 + NW - 1) / NW, IB = NI * NW - IA;
    constexpr int SBP = IB * 128;                           // B stage with the spare slots
    static_assert(SA % 128 == 0 && SB % 128 == 0, "stage size");
    static_assert(NI < 16, "load count per wave");
    static_assert(TM % 16 == 0 && TN % 16 == 0 && BM % WM == 0 && BN % WN == 0, "wave tile");
    static_assert(BM % 32 == 0, "A pitch 16 (mod 32)");
    if (S && S->err) return;
    // the three stages in one block, which the fused epilogue of the NN form
    // reuses for the tile (BN rows of pitch BM + 1)
    constexpr int SMN = 3 * (SA + SBP), STN = NN ? BN * (BM + 1) : 0;
    __shared__ __attribute__((aligned(1024))) double SM[SMN > STN ? SMN : STN];
    double* const S0 = SM;
    double* const S1 = SM + (SA + SBP);
    double* const S2 = SM + 2 * (SA + SBP);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    int bx = vb.x, by = vb.y;
    constexpr int PM = GEMM_PATCH_M, PN = 32 / GEMM_PATCH_M;
    if (vb.gx % PM == 0 && vb.gy % PN == 0 && (vb.gx * vb.gy) % 256 == 0) {
        // dispatch id -> XCD (id % 8); 32 consecutive ids of one XCD form a
        // PM x PN patch (a bijection when every XCD receives whole patches)
        const int id = vb.x + vb.gx * vb.y, x = id % 8, l = id / 8;
        const int npm = vb.gx / PM;
        const int patch = x + 8 * (l / 32), ll = l % 32;
        bx = (patch % npm) * PM + ll % PM;
        by = (patch / npm) * PN + ll / PM;
    }
    const int i0 = bx * BM, n0 = by * BN;
    int kb, ke;
    tile_krange(ks, K, i0, BM, BK, kb, ke);
    const int mlast = M - 2 + (M & 1), klast = K - 2 + (K & 1), nlast = N - 2 + (N & 1);

    auto issue = [&](int k0, double* As, double* Bs) {
#pragma unroll
        for (int r = 0; r < NI; ++r) {
            const int ii = wave + NW * r;
            const double* src;
            double* dst;
            if (ii < IA) {
                const int slot = ii * 64 + lane, kk = slot / SLA, q = slot % SLA;
                src = A + (size_t)min(k0 + kk, K - 1) * lda + min(i0 + 2 * q, mlast);
                dst = As + ii * 128;
            } else {
                const int jj = ii - IA, slot = jj * 64 + lane, row = slot / SLB, q = slot % SLB;
                // rows past the tile (the spare slots) read clamped addresses
                if (NN) src = B + (size_t)min(n0 + row, N - 1) * ldb + min(k0 + 2 * q, klast);
                else src = B + (size_t)min(k0 + min(row, BK - 1), K - 1) * ldb + min(n0 + 2 * q, nlast);
                dst = Bs + jj * 128;
            }
            dma16(src, dst);
        }
    };

    f64x4 acc[RM][RN];
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < RN; ++c) acc[r][c] = (f64x4){0.0, 0.0, 0.0, 0.0};

    const int nk = (ke - kb + BK - 1) / BK;
    // slice s is in stage (As, Bs); slice s+2 goes into (An, Bn), the stage of slice s-1
    auto step = [&](int s, double* As, double* Bs, double* An, double* Bn) {
        if (s + 1 < nk) __builtin_amdgcn_s_waitcnt(0x0F70 | NI);   // vmcnt(NI): only slice s+1 in flight
        else __builtin_amdgcn_s_waitcnt(0x0F70);                    // vmcnt(0)
        raw_barrier();
        const int kv = ke - kb - s * BK;      // valid k-rows of this slice
        if (kv < BK) {
            for (int e = t; e < (BK - kv) * LA; e += 64 * NW) As[kv * LA + e] = 0.0;
            if (NN) {
                for (int e = t; e < BN * (BK - kv); e += 64 * NW) Bs[(e / (BK - kv)) * LB + kv + e % (BK - kv)] = 0.0;
            } else {
                for (int e = t; e < (BK - kv) * LB; e += 64 * NW) Bs[kv * LB + e] = 0.0;
            }
            __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
            raw_barrier();
        }
#if GEMM_PRIO == 2
        __builtin_amdgcn_s_setprio(1);
#endif
        if (s + 2 < nk) issue(kb + (s + 2) * BK, An, Bn);
#if GEMM_PRIO == 2
        __builtin_amdgcn_s_setprio(0);
#endif
#if GEMM_PRIO == 1
        __builtin_amdgcn_s_setprio(1);
#endif
#pragma unroll
        for (int k4 = 0; k4 < BK; k4 += 4) {
            const int kr = k4 + (lane >> 4);
            double a[RM], b[RN];
#pragma unroll
            for (int r = 0; r < RM; ++r) a[r] = As[kr * LA + wm * TM + r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < RN; ++c)
                b[c] = NN ? Bs[(wn * TN + c * 16 + (lane & 15)) * LB + kr] : Bs[kr * LB + wn * TN + c * 16 + (lane & 15)];
#pragma unroll
            for (int r = 0; r < RM; ++r)
#pragma unroll
                for (int c = 0; c < RN; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
#if GEMM_PRIO == 1
        __builtin_amdgcn_s_setprio(0);
#endif
    };
    issue(kb, S0, S0 + SA);
    if (nk > 1) issue(kb + BK, S1, S1 + SA);
    for (int s = 0; s < nk; s += 3) {
        step(s, S0, S0 + SA, S2, S2 + SA);
        if (s + 1 < nk) step(s + 1, S1, S1 + SA, S0, S0 + SA);
        if (s + 2 < nk) step(s + 2, S2, S2 + SA, S1, S1 + SA);
    }
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < RN; ++c) {
            const int col = n0 + wn * TN + c * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = i0 + wm * TM + r * 16 + (lane >> 4) + 4 * v;
                if (row < M && col < N) C[(size_t)col * ldc + row] = acc[r][c][v];
            }
        }
    if (zg) {
        __shared__ double gsh[NW];
        gauge_partials<RM, RN, TM, TN, NW, BM>(acc, ks, M, N, i0, n0, wm, wn, lane, wave, by, vb.gy, zg, gpart, gsh);
    }
    if (NN && yf.csz > 0) {
        // chunk elimination of the y-mode z-lines on this tile: rows i are
        // y-modes, columns n the z rows; the tile holds BN / csz whole chunks
        constexpr int LT = BM + 1;   // LDS row pitch of the staged tile (one z row per LDS row)
        lds_barrier();   // no DMA is in flight after the last slice
#pragma unroll
        for (int r = 0; r < RM; ++r)
#pragma unroll
            for (int c = 0; c < RN; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    SM[(wn * TN + c * 16 + (lane & 15)) * LT + wm * TM + r * 16 + (lane >> 4) + 4 * v] = acc[r][c][v];
        lds_barrier();
        constexpr int WPC = (BM + 63) / 64;   // waves per chunk
        const int csz = yf.csz, cc = wave / WPC, iw = (wave % WPC) * 64 + lane;
        if (cc < BN / csz) {                  // wave-uniform
            const int il = iw < BM ? iw : BM - 1;
            const int ic = min(i0 + il, M - 1);
            const int p = (n0 + cc * csz) / csz, s = p * csz, m = csz - 1;
            const double* col = SM + (size_t)(cc * csz) * LT + il;
            double dv[32];
#pragma unroll
            for (int q = 0; q < 32; ++q) dv[q] = col[(q < m ? q : 0) * LT];
            const double w = yf.wy[ic];
            double yl, vl, ys, vs, ws, wl;
            ychunk_rows<32, false>(yf.amphk, yf.acphk, yf.apphk, s, m, w, dv, nullptr, nullptr, nullptr, yl, vl,
                                   ys, vs, ws, wl);
            if (iw < BM && i0 + iw < M)
                ysum_store(yf.ysum, yf.n2m, i0 + iw + 1, p, yf.amphk, yf.acphk, s + m, w, col[m * LT], yl, vl, wl,
                           ys, vs, ws);
        }
    }
}
template <int BK, bool NN, int WM = 2, int WN = 8, int BN = 128, int BM = 96>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_dma(int M, int N, int K, const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc, const DevScalars* S, const double* __restrict__ zg, double* gpart, YFuse yf, int ks)
{
    gemm_dma_body<BK, NN, WM, WN, BN, BM>(M, N, K, A, lda, B, ldb, C, ldc, S, zg, gpart, yf, ks, vb_hw());
}
template <int BK, bool NN, int WM = 2, int WN = 8, int BN = 128, int BM = 96>
__global__ __launch_bounds__(64 * WM * WN) void kb_gemm_dma(const BGrid<GemmProb> G)
{
    VBlock vb;
    const GemmProb& P = *bg_block(G, vb);
    gemm_dma_body<BK, NN, WM, WN, BN, BM>(P.M, P.N, P.K, P.A, P.lda, P.B, P.ldb, P.C, P.ldc, P.S, P.zg, P.gpart, P.yf, P.ks, vb);
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// The direct-to-LDS path reads operand pairs (16 bytes): both operands must be
// 16-byte aligned with even leading dimensions, and a pair that starts on the
// last row of an odd extent reads one padding element, which must exist.
template <bool NN>
static bool dma_ok(int M, int N, int K, const double* A, int lda, const double* B, int ldb)
{
    const int bext = NN ? K : N;
    return ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0 && lda % 2 == 0 &&
           ldb % 2 == 0 && lda >= M + (M & 1) && ldb >= bext + (bext & 1);
}

// One launch; with events, the dispatch itself records its start and end
// (hipExtLaunchKernel), so timing adds no packets between kernels.
template <class Kern, class... Args>
static void launch_timed(Kern kern, dim3 g, dim3 b, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0) hipExtLaunchKernelGGL(kern, g, b, 0, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kern, g, b, 0, st, args...);
}

template <bool NN>
static int launch_gemm(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                       const DevScalars* S, hipStream_t st, const double* zg, double* gpart, bool plan,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const YFuse* yfp = nullptr,
                       bool* fused = nullptr, int ks = 0)
{
    YFuse yf{};
    // the fused chunk pass needs whole chunks of at most 33 rows in every
    // output tile of the direct-to-LDS kernel, and no more chunks per tile
    // than the epilogue has wave groups (one group of (96 + 63) / 64 waves
    // per chunk)
    constexpr int kChunkGroups = (GEMM_DMA_WM * GEMM_DMA_WN) / ((96 + 63) / 64);
    if (NN && yfp && yfp->csz > 1 && yfp->csz <= 33 && GEMM_DMA_BN % yfp->csz == 0 && N % GEMM_DMA_BN == 0 &&
        GEMM_DMA_BN / yfp->csz <= kChunkGroups)
        yf = *yfp;
    if (fused) *fused = false;
    if ((long)cdiv(M, 96) * cdiv(N, 128) >= 192 && dma_ok<NN>(M, N, K, A, lda, B, ldb)) {
        if (!plan)
            launch_timed(k_gemm_dma<GEMM_DMA_BK, NN, GEMM_DMA_WM, GEMM_DMA_WN, GEMM_DMA_BN>,
                         dim3(cdiv(M, 96), cdiv(N, GEMM_DMA_BN)), dim3(64 * GEMM_DMA_WM * GEMM_DMA_WN), st, e0, e1, M,
                         N, K, A, lda, B, ldb, C, ldc, S, zg, gpart, yf, ks);
        if (fused) *fused = yf.csz > 0;
        return cdiv(N, GEMM_DMA_BN);
    } else if ((long)cdiv(M, 96) * cdiv(N, 128) >= 192) {
        if (!plan)
            launch_timed(k_gemm_nt<96, 128, 2, 4, NN>, dim3(cdiv(M, 96), cdiv(N, 128)), dim3(512), st, e0, e1, M, N,
                         K, A, lda, B, ldb, C, ldc, S, zg, gpart, ks);
        return cdiv(N, 128);
    } else if ((long)cdiv(M, 64) * cdiv(N, 64) >= 96) {
        if (!plan)
            launch_timed(k_gemm_nt<64, 64, 2, 2, NN>, dim3(cdiv(M, 64), cdiv(N, 64)), dim3(256), st, e0, e1, M, N,
                         K, A, lda, B, ldb, C, ldc, S, zg, gpart, ks);
        return cdiv(N, 64);
    }
    if (!plan)
        launch_timed(k_gemm_nt<32, 32, 2, 2, NN>, dim3(cdiv(M, 32), cdiv(N, 32)), dim3(256), st, e0, e1, M, N, K,
                     A, lda, B, ldb, C, ldc, S, zg, gpart, ks);
    return cdiv(N, 32);
}

void launch_gemm_nt(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                    const DevScalars* S, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    launch_gemm<false>(M, N, K, A, lda, B, ldb, C, ldc, S, st, nullptr, nullptr, false, e0, e1);
}

bool launch_gemm_nn(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                    const DevScalars* S, hipStream_t st, const double* zg, double* gpart, hipEvent_t e0,
                    hipEvent_t e1, const YFuse* yf, int ks)
{
    bool fused = false;
    launch_gemm<true>(M, N, K, A, lda, B, ldb, C, ldc, S, st, zg, gpart, false, e0, e1, yf, &fused, ks);
    return fused;
}

// Tile columns of the C = A * B launch these arguments get (the number of
// gauge partials it writes).
int gemm_nn_tiles_n(int M, int N, int K, const double* A, int lda, const double* B, int ldb)
{
    return launch_gemm<true>(M, N, K, A, lda, B, ldb, nullptr, 0, nullptr, 0, nullptr, nullptr, true);
}

// ---------------------------------------------------------------------------
// Batched transforms of the Poisson solve: one grouped launch over the
// instances' problems (each instance's tiles are its own grid), one tile
// shape for the whole batch (B.gemm_cfg).  Every tile shape accumulates each
// output element over k in the same order, so the transform results do not
// depend on the shape; the gauge partials are one per tile column of the
// shape in use (gemm_cfg_tile_n).
// Shapes 3 and 4 are the direct-to-LDS kernel with 96 x GEMM_DMA_BN tiles of
// 16 waves (the single-instance transforms' shape, K-slices of GEMM_DMA_BK)
// and 96 x 64 tiles of 8 waves, two workgroups per CU, with K-slices of 16
// (faster when many small instances leave partial tiles and short K ranges:
// on the 256-member ensemble 1.56 ms per grouped transform against 1.88
// with 24-deep slices).
int gemm_cfg_tile_n(int cfg)
{
    return cfg == 0 ? 32 : cfg == 1 ? 64 : cfg == 2 ? 128 : cfg == 3 ? GEMM_DMA_BN : 64;
}

// The shapes every instance of the batch can run: bit c set for cfg c.
int gemm_cfg_mask(const FluidDev& F)
{
    int m = 7;
    const bool yax = F.taxis == 1;
    const double* A = yax ? F.ymmi : F.dph + fidx(F, 1, 1);
    const double* Bm = yax ? F.dph + fidx(F, 1, 1) : F.zmmt;
    const int lda = yax ? F.ldy : F.P, ldb = yax ? F.P : F.ldz;
    const int K = yax ? F.n2m : F.n3m;
    const bool ok1 = yax ? dma_ok<true>(F.n2m, F.n3m, K, A, lda, Bm, ldb) : dma_ok<false>(F.n2m, F.n3m, K, A, lda, Bm, ldb);
    const double* A2 = yax ? F.ymm : F.hmn;
    const double* B2 = yax ? F.hmn : F.zmm;
    const bool ok2 = yax ? dma_ok<true>(F.n2m, F.n3m, K, A2, lda, B2, ldb) : dma_ok<false>(F.n2m, F.n3m, K, A2, F.P, B2, ldb);
    if (ok1 && ok2) m |= 8 | 16;
    return m;
}

// Tile columns (gauge partials) of the backward transform with shape cfg.
int gemm_cfg_parts(const FluidDev& F, int cfg) { return cdiv(F.n3m, gemm_cfg_tile_n(cfg)); }

template <bool NN>
static void launch_gemm_b(BatchSet& B, int key, const std::vector<GemmProb>& pr, hipStream_t st)
{
    std::vector<int> d(3 * (size_t)B.n);
    int bm = 32, bn = 32;
    switch (B.gemm_cfg) {
    case 1: bm = 64; bn = 64; break;
    case 2: bm = 96; bn = 128; break;
    case 3: bm = 96; bn = GEMM_DMA_BN; break;
    case 4: bm = 96; bn = 64; break;
    default: break;
    }
    for (int i = 0; i < B.n; ++i) {
        d[3 * i] = cdiv(pr[i].M, bm);
        d[3 * i + 1] = cdiv(pr[i].N, bn);
        d[3 * i + 2] = 1;
    }
    const int tot = bg_total(d);
    if (tot <= 0) return;
    const BGrid<GemmProb> G = B.gemm(key, pr, d);
    switch (B.gemm_cfg) {
    case 3:
        hipLaunchKernelGGL((kb_gemm_dma<GEMM_DMA_BK, NN, GEMM_DMA_WM, GEMM_DMA_WN, GEMM_DMA_BN>), dim3(tot),
                           dim3(64 * GEMM_DMA_WM * GEMM_DMA_WN), 0, st, G);
        break;
    case 4: hipLaunchKernelGGL((kb_gemm_dma<16, NN, 2, 4, 64>), dim3(tot), dim3(512), 0, st, G); break;
    case 2: hipLaunchKernelGGL((kb_gemm_nt<96, 128, 2, 4, NN>), dim3(tot), dim3(512), 0, st, G); break;
    case 1: hipLaunchKernelGGL((kb_gemm_nt<64, 64, 2, 2, NN>), dim3(tot), dim3(256), 0, st, G); break;
    default: hipLaunchKernelGGL((kb_gemm_nt<32, 32, 2, 2, NN>), dim3(tot), dim3(256), 0, st, G); break;
    }
}

bool launch_poisson_gemm_b(BatchSet& B, bool fwd, hipStream_t st)
{
    const bool yax = B.hF[0].taxis == 1;
    std::vector<GemmProb> pr(B.n);
    bool fused = false;
    for (int i = 0; i < B.n; ++i) {
        const FluidDev& F = B.hF[i];
        GemmProb& q = pr[i];
        q = GemmProb{};
        q.M = F.n2m;
        q.N = F.n3m;
        q.S = F.S;
        double* d = F.dph + fidx(F, 1, 1);
        if (yax) {
            q.K = F.n2m;
            q.lda = F.ldy;
            q.ldb = F.P;
            q.ldc = F.P;
            q.A = F.ysh > 0 ? (fwd ? F.ysmi : F.ysm) : (fwd ? F.ymmi : F.ymm);
            q.ks = F.ysh;
            q.B = fwd ? d : F.hmn;
            q.C = fwd ? F.hmn : d;
            if (!fwd) {
                q.zg = F.zg;
                q.gpart = F.gpart;
            } else if (B.ysolve_fused) {
                q.yf = YFuse{F.wy, F.amphk, F.acphk, F.apphk, F.ysum, F.n2m, F.n3m / F.ks[0].nch};
                fused = true;
            }
        } else {
            q.K = F.n3m;
            q.lda = F.P;
            q.ldb = F.ldz;
            q.ldc = F.P;
            q.A = fwd ? d : F.hmn;
            q.B = fwd ? F.zmmt : F.zmm;
            q.C = fwd ? F.hmn : d;
        }
    }
    if (yax) launch_gemm_b<true>(B, fwd ? BK_GEMM_F : BK_GEMM_B, pr, st);
    else launch_gemm_b<false>(B, fwd ? BK_GEMM_F : BK_GEMM_B, pr, st);
    return fused;
}

// The fused chunk pass is possible for an instance when its line partition
// has equal chunks that fill the direct-to-LDS kernel's output tiles of batch
// shape cfg (3 or 4), and the tile has a group of (96 + 63) / 64 waves per
// chunk.
bool gemm_fuse_ok(const FluidDev& F, int cfg)
{
    if (cfg != 3 && cfg != 4) return false;
    const int bn = gemm_cfg_tile_n(cfg), waves = cfg == 3 ? GEMM_DMA_WM * GEMM_DMA_WN : 8;
    const int groups = waves / ((96 + 63) / 64);
    if (F.taxis != 1 || F.n3m % F.ks[0].nch != 0) return false;
    const int csz = F.n3m / F.ks[0].nch;
    return csz > 1 && csz <= 33 && bn % csz == 0 && F.n3m % bn == 0 && bn / csz <= groups;
}

// ---------------------------------------------------------------------------
// FP32 transforms of the mixed-precision Poisson solve: C = X0 + A * B with
// the operands read in FP64 (the FP64 eigen-matrices and fields), rounded to
// FP32 as they are staged in LDS, multiplied on v_mfma_f32_16x16x4_f32 with
// FP32 accumulation, and the result written back in FP64 (added to X0 when
// X0 is given).  Only the C = A * B form of the transforms along y ("NN": B
// is k-contiguous).  Fragment layout of v_mfma_f32_16x16x4_f32: A lane l
// holds A[l%16][l/16], B lane l holds B[l/16][l%16]; accumulator element v
// of lane l is C[4*(l/16) + v][l%16].  The gauge partials (zg) are taken
// from the final FP64 values of row M-1.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm32(int M, int N, int K, const double* __restrict__ A, int lda,
                                                         const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                         int ldc, const double* __restrict__ X0,
                                                         const DevScalars* S, const double* __restrict__ zg,
                                                         double* gpart)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int RM = TM / 16, RN = TN / 16;
    constexpr int LA = BM + 4;                   // A: k-rows of BM floats
    constexpr int LB = BN + 1;                   // B: k-rows of BN floats (transposing store)
    constexpr int NA = (BM * GBK + NT - 1) / NT;
    constexpr int NB = (BN * GBK + NT - 1) / NT;
    static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile");
    if (S && S->err) return;
    __shared__ float As[2][GBK][LA];
    __shared__ float Bs[2][GBK][LB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int i0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    float ra[NA], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = t + u * NT, kk = e / BM, r = e % BM;
            const bool ok = (e < BM * GBK) && (k0 + kk < K) && (i0 + r < M);
            ra[u] = ok ? (float)A[(size_t)(k0 + kk) * lda + i0 + r] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = t + u * NT, kk = e % GBK, r = e / GBK;
            const bool ok = (e < BN * GBK) && (k0 + kk < K) && (n0 + r < N);
            rb[u] = ok ? (float)B[(size_t)(n0 + r) * ldb + k0 + kk] : 0.0f;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = t + u * NT;
            if (e < BM * GBK) As[buf][e / BM][e % BM] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = t + u * NT;
            if (e < BN * GBK) Bs[buf][e % GBK][e / GBK] = rb[u];
        }
    };
    f32x4 acc[RM][RN];
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < RN; ++c) acc[r][c] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += GBK) {
        const bool more = k0 + GBK < K;
        if (more) load(k0 + GBK);
#pragma unroll
        for (int k4 = 0; k4 < GBK; k4 += 4) {
            const int kr = k4 + (lane >> 4);
            float a[RM], b[RN];
#pragma unroll
            for (int r = 0; r < RM; ++r) a[r] = As[buf][kr][wm * TM + r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < RN; ++c) b[c] = Bs[buf][kr][wn * TN + c * 16 + (lane & 15)];
#pragma unroll
            for (int r = 0; r < RM; ++r)
#pragma unroll
                for (int c = 0; c < RN; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    double g = 0.0;
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < RN; ++c) {
            const int col = n0 + wn * TN + c * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = i0 + wm * TM + r * 16 + 4 * (lane >> 4) + v;
                if (row < M && col < N) {
                    const size_t q = (size_t)col * ldc + row;
                    const double x = (X0 ? X0[q] : 0.0) + (double)acc[r][c][v];
                    C[q] = x;
                    if (zg && row == M - 1) g = g + zg[col] * x;
                }
            }
        }
    if (zg && M - 1 >= i0 && M - 1 < i0 + BM) {
        __shared__ double gsh[WM * WN];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) g = g + __shfl_xor(g, o, 64);
        if (lane == 0) gsh[wave] = g;
        __syncthreads();
        if (t == 0) {
            double s = gsh[0];
            for (int w = 1; w < WM * WN; ++w) s = s + gsh[w];
            gpart[blockIdx.y] = s;
        }
    }
}

// Tile columns of the FP32 transform (the gauge partials it writes).
int gemm32_tiles_n(int N) { return cdiv(N, 128); }

void launch_gemm32(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                   const double* X0, const DevScalars* S, hipStream_t st, const double* zg, double* gpart,
                   hipEvent_t e0, hipEvent_t e1)
{
    launch_timed(k_gemm32<128, 128, 2, 4>, dim3(cdiv(M, 128), cdiv(N, 128)), dim3(512), st, e0, e1, M, N, K, A, lda,
                 B, ldb, C, ldc, X0, S, zg, gpart);
}

// Diagnostic entry point: the FP32 transform on host arrays, C = A * B
// (column-major, lda/ldb/ldc), for the layout and accuracy test.
extern "C" int fribo_test_gemm32(int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                                 int32_t ldb, double* C, int32_t ldc)
{
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || lda < M || ldb < K || ldc < M) return 5;
    double *a = nullptr, *b = nullptr, *c = nullptr;
    const size_t na = (size_t)lda * K, nb = (size_t)ldb * N, nc = (size_t)ldc * N;
    if (hipMalloc(&a, na * 8) != hipSuccess || hipMalloc(&b, nb * 8) != hipSuccess ||
        hipMalloc(&c, nc * 8) != hipSuccess) {
        hipFree(a);
        hipFree(b);
        hipFree(c);
        return 6;
    }
    int rc = 0;
    if (hipMemcpy(a, A, na * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b, B, nb * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c, C, nc * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = 6;
    if (!rc) {
        launch_gemm32(M, N, K, a, lda, b, ldb, c, ldc, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(C, c, nc * 8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = 6;
    }
    hipFree(a);
    hipFree(b);
    hipFree(c);
    return rc;
}
