// FP64 MFMA GEMM for the Poisson eigen-transforms (phcalc_outflow.F90:31,61).
//
//   C[i + n*ldc] = sum_kk A[i + kk*lda] * B[n + kk*ldb]     (C = A * B^T)
//
// Both transforms have this form on the device: the forward transform
// hmn = dph * zmmt^T takes A = dph (j contiguous) and B = zmmt; the backward
// transform dph = V * zmm^T takes A = V and B = zmm in the reference's own
// column-major layout.  Every K-slice of both operands is therefore a set of
// contiguous columns and loads fully coalesced.
//
// Tiling: 64x64 output tile per 256-thread workgroup (2x2 waves of 32x32),
// BK = 16, v_mfma_f64_16x16x4_f64 (A fragment: lane l holds A[l%16][l/16];
// B fragment: B[l/16][l%16]; accumulator element v of lane l is
// C[l/16 + 4v][l%16]).  The next K-slice is fetched into registers while the
// current one is consumed from LDS.
#include "fribo_internal.h"

typedef double f64x4 __attribute__((ext_vector_type(4)));

#define GBM 64
#define GBN 64
#define GBK 16
#define GPAD 16

__global__ __launch_bounds__(256) void k_gemm_nt(int M, int N, int K, const double* __restrict__ A, int lda,
                                                 const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                 int ldc, const DevScalars* S)
{
    if (S && S->err) return;
    __shared__ double As[GBK][GBM + GPAD];
    __shared__ double Bs[GBK][GBN + GPAD];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int i0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
    // loader mapping: column kk = t/16, rows (t%16)*4 .. +3
    const int lk = t >> 4, lr = (t & 15) * 4;
    double ra[4], rb[4];

    auto load = [&](int k0) {
        const int kk = k0 + lk;
        const bool kin = kk < K;
        const double* pa = A + (size_t)kk * lda + i0 + lr;
        const double* pb = B + (size_t)kk * ldb + n0 + lr;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ra[v] = (kin && i0 + lr + v < M) ? pa[v] : 0.0;
            rb[v] = (kin && n0 + lr + v < N) ? pb[v] : 0.0;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            As[lk][lr + v] = ra[v];
            Bs[lk][lr + v] = rb[v];
        }
    };

    f64x4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = (f64x4){0.0, 0.0, 0.0, 0.0};

    load(0);
    store();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += GBK) {
        const bool more = k0 + GBK < K;
        if (more) load(k0 + GBK);
#pragma unroll
        for (int k4 = 0; k4 < GBK; k4 += 4) {
            const int kr = k4 + (lane >> 4);
            double a0 = As[kr][wm * 32 + (lane & 15)];
            double a1 = As[kr][wm * 32 + 16 + (lane & 15)];
            double b0 = Bs[kr][wn * 32 + (lane & 15)];
            double b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            store();
            __syncthreads();
        }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int col = n0 + wn * 32 + c * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = i0 + wm * 32 + r * 16 + (lane >> 4) + 4 * v;
                if (row < M && col < N) C[(size_t)col * ldc + row] = acc[r][c][v];
            }
        }
}

void launch_gemm_nt(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                    const DevScalars* S, hipStream_t st)
{
    dim3 g((M + GBM - 1) / GBM, (N + GBN - 1) / GBN);
    hipLaunchKernelGGL(k_gemm_nt, g, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, S);
}
