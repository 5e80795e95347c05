// Poisson solve with the eigen-transform along y (used when n2m < n3m).
//
// phcalc_outflow.F90:4-71 solves (A_y (x) I + I (x) A_z) dph = rhs by
// diagonalising A_z (zmm, zmmt, wr) and solving one tridiagonal system along
// y per z-mode.  Diagonalising A_y instead (Y, Yinv, wy; fribo_eig.hip)
// gives the same discrete solution with one tridiagonal system along z per
// y-mode and a dense transform of 2*n3m*n2m^2 flops instead of
// 2*n2m*n3m^2:
//
//   hmn = Yinv * dph           (GEMM, B operand k-contiguous)
//   (A_z + wy(j)) v(j,:) = hmn(j,:)   for every y-mode j   (this file)
//   dph = Y * v                (GEMM)
//   dph -= c                   (gauge: partials from the second GEMM, subtracted by k_upd)
//
// The Neumann problem is singular (constants); the reference's order pins
// the z null mode's last unknown, which fixes the constant through
// sum_k zmmt(k0,k) dph(n2m,k) = 0.  Here the y null mode is pinned instead;
// the epilogue of the second GEMM leaves the partial sums of the constant c
// that restores the reference's gauge, and k_upd subtracts c when it reads
// dph (S->dphoff keeps c for the read-back of dph).
//
// The per-mode z-lines use the partition of the q2 k-system (ks[0]: n3m rows,
// up to 64 chunks).  Lanes run along the y-mode index (coalesced), each
// thread owns one chunk of one line; the matrix rows amphk, acphk + wy(j),
// apphk are read as wave-uniform scalars, the shift per lane.  Three passes
// as in the Helmholtz k-solves: chunk elimination -> separator system (one
// wave per line, PCR) -> chunk back-substitution.
#include "fribo_ychunk.h"
#include "fribo_batch.h"

// Waves per workgroup (each wave its own 64 y-modes, or its own mode in the
// separator pass).
#ifndef FRIBO_YW
#define FRIBO_YW 1
#endif
#define YW FRIBO_YW

__device__ inline size_t yidx(const FluidDev& F, int q, int j, int p) { return ysum_idx(F.n2m, q, j, p); }

// Chunk elimination of line j, chunk p (rows s..e-1 interior, e separator),
// the rows read from column col of hmn.
template <int MAXM, bool STORE>
__device__ inline void ychunk(const FluidDev& F, const KSys& K, int p, double w, const double* col, double* cp,
                              double* vp, double* y, double& yl, double& vl, double& ys, double& vs, double& ws,
                              double& wl)
{
    const size_t P = F.P;
    const int s = K.cs[p], e = K.cs[p + 1] - 1, m = e - s;
    double dv[MAXM];
#pragma unroll
    for (int t = 0; t < MAXM; ++t) dv[t] = col[(size_t)(s + (t < m ? t : 0)) * P];
    ychunk_rows<MAXM, STORE>(F.amphk, F.acphk, F.apphk, s, m, w, dv, cp, vp, y, yl, vl, ys, vs, ws, wl);
}

// Y1C chunks of the same 64 y-modes per workgroup, one wave per chunk
// (lanes past the last y-mode work on mode n2m and store nothing: every lane
// stays active through the row broadcasts).  The summaries are chunk-fastest
// in memory, so they are transposed in LDS and stored as runs of Y1C chunks
// per mode instead of one 8-byte store per lane.
// (16 chunks per workgroup for chunks of up to 32 rows; 4 for longer ones,
// whose registers allow fewer waves per SIMD)
__host__ __device__ constexpr int y1c(int maxm) { return maxm <= 32 ? 16 : 4; }
template <int MAXM>
__device__ __forceinline__ void ysolve1_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int Y1C = y1c(MAXM);
    __shared__ double sy[6][Y1C][65];
    const DevScalars* S = F.S;
    if (S->err) return;
    const KSys& K = F.ks[0];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = vb.y * Y1C + wave;
    const int j0 = vb.x * 64;
    const int jc = min(j0 + lane + 1, F.n2m);
    if (p < K.nch) {
        const double* col = F.hmn + (jc - 1);
        const double w = F.wy[jc - 1];
        double yl, vl, ys, vs, ws, wl;
        ychunk<MAXM, false>(F, K, p, w, col, nullptr, nullptr, nullptr, yl, vl, ys, vs, ws, wl);
        const int e = K.cs[p + 1] - 1;
        double v[6];
        ysum_vals(v, F.amphk, F.acphk, e, w, col[(size_t)e * F.P], yl, vl, wl, ys, vs, ws);
#pragma unroll
        for (int q = 0; q < 6; ++q) sy[q][wave][lane] = v[q];
    }
    __syncthreads();
    const int mode = threadIdx.x / Y1C, c = threadIdx.x % Y1C;
    const int pj = vb.y * Y1C + c, jj = j0 + mode + 1;
    if (pj < K.nch && jj <= F.n2m)
#pragma unroll
        for (int q = 0; q < 6; ++q) F.ysum[yidx(F, q, jj, pj)] = sy[q][c][mode];
}
template <int MAXM>
__global__ __launch_bounds__(64 * y1c(MAXM)) void k_ysolve1(const FluidDev F)
{
    ysolve1_body<MAXM>(F, vb_hw());
}
template <int MAXM>
__global__ __launch_bounds__(64 * y1c(MAXM)) void kb_ysolve1(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ysolve1_body<MAXM>(*Fp, vb);
}

__device__ __forceinline__ void ysolve2_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const KSys& K = F.ks[0];
    const int j = vb.x * YW + (threadIdx.x >> 6) + 1, p = threadIdx.x & 63, lane = p;
    if (j > F.n2m) return;   // whole wave
    const int nch = K.nch;
    const double* Y = F.ysum;
    double alpha = 0.0, beta = 1.0, gamma = 0.0, delta = 0.0;
    if (p < nch) {
        alpha = Y[yidx(F, 0, j, p)];
        beta = Y[yidx(F, 1, j, p)];
        delta = Y[yidx(F, 2, j, p)];
        if (p + 1 < nch) {
            const double ce = F.apphk[K.cs[p + 1]];   // row e = cs[p+1]-1, k = e+1
            beta = beta + ce * Y[yidx(F, 4, j, p + 1)];
            gamma = ce * Y[yidx(F, 5, j, p + 1)];
            delta = delta - ce * Y[yidx(F, 3, j, p + 1)];
        }
        if (j == F.ynull && p == nch - 1) {
            alpha = 0.0;
            beta = 1.0;
            gamma = 0.0;
            delta = 0.0;
        }
    }
    for (int h = 1; h < 64; h <<= 1) {
        double am = __shfl_up(alpha, h, 64), bm = __shfl_up(beta, h, 64);
        double gm = __shfl_up(gamma, h, 64), dm = __shfl_up(delta, h, 64);
        double ap = __shfl_down(alpha, h, 64), bp = __shfl_down(beta, h, 64);
        double gp = __shfl_down(gamma, h, 64), dp = __shfl_down(delta, h, 64);
        if (lane - h < 0) { am = 0.0; bm = 1.0; gm = 0.0; dm = 0.0; }
        if (lane + h >= 64) { ap = 0.0; bp = 1.0; gp = 0.0; dp = 0.0; }
        const double m1 = alpha * recip(bm), m2 = gamma * recip(bp);
        const double na = -m1 * am, ng = -m2 * gp;
        const double nb = beta - m1 * gm - m2 * ap;
        const double nd = delta - m1 * dm - m2 * dp;
        alpha = na;
        beta = nb;
        gamma = ng;
        delta = nd;
    }
    if (p < nch) F.ysum[yidx(F, 6, j, p)] = delta * recip(beta);
}
__global__ __launch_bounds__(64 * YW) void k_ysolve2(const FluidDev F)
{
    ysolve2_body(F, vb_hw());
}
__global__ __launch_bounds__(64 * YW) void kb_ysolve2(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ysolve2_body(*Fp, vb);
}

template <int MAXM>
__device__ __forceinline__ void ysolve3_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const KSys& K = F.ks[0];
    const int p = vb.y;
    if (p >= K.nch) return;
    const int j = (vb.x * YW + (threadIdx.x >> 6)) * 64 + (threadIdx.x & 63) + 1;
    const bool act = j <= F.n2m;
    const int jc = act ? j : F.n2m;
    double* col = F.hmn + (jc - 1);
    const double w = F.wy[jc - 1];
    double cp[MAXM], vp[MAXM], y[MAXM];
    double yl, vl, ys, vs, ws, wl;
    ychunk<MAXM, true>(F, K, p, w, col, cp, vp, y, yl, vl, ys, vs, ws, wl);
    if (!act) return;
    const double X = F.ysum[yidx(F, 6, j, p)];
    const double Xm = p > 0 ? F.ysum[yidx(F, 6, j, p - 1)] : 0.0;
    const size_t P = F.P;
    const int s = K.cs[p], e = K.cs[p + 1] - 1, m = e - s;
    double x = 0.0;
#pragma unroll
    for (int t = MAXM - 1; t >= 0; --t) {
        if (t < m) {
            if (t == m - 1) x = y[t] + vp[t] * Xm + wl * X;
            else x = (y[t] + vp[t] * Xm) - cp[t] * x;
            col[(size_t)(s + t) * P] = x;
        }
    }
    col[(size_t)e * P] = X;
}
template <int MAXM>
__global__ __launch_bounds__(64 * YW) void k_ysolve3(const FluidDev F)
{
    ysolve3_body<MAXM>(F, vb_hw());
}
template <int MAXM>
__global__ __launch_bounds__(64 * YW) void kb_ysolve3(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ysolve3_body<MAXM>(*Fp, vb);
}

void launch_ysolve(const FluidDev& F, hipStream_t st, bool first_done)
{
    const int mm = F.ks[0].maxm - 1;   // interior rows per chunk
    dim3 g((F.n2m + 64 * YW - 1) / (64 * YW), F.ks[0].nch), b(64 * YW);
    if (!first_done) {
        const int c = y1c(mm <= 32 ? 32 : 64);
        const dim3 g1((F.n2m + 63) / 64, (F.ks[0].nch + c - 1) / c), b1(64 * c);
        if (mm <= 16) hipLaunchKernelGGL(k_ysolve1<16>, g1, b1, 0, st, F);
        else if (mm <= 32) hipLaunchKernelGGL(k_ysolve1<32>, g1, b1, 0, st, F);
        else if (mm <= 64) hipLaunchKernelGGL(k_ysolve1<64>, g1, b1, 0, st, F);
        else hipLaunchKernelGGL(k_ysolve1<128>, g1, b1, 0, st, F);
    }
    hipLaunchKernelGGL(k_ysolve2, dim3((F.n2m + YW - 1) / YW), b, 0, st, F);
    if (mm <= 16) hipLaunchKernelGGL(k_ysolve3<16>, g, b, 0, st, F);
    else if (mm <= 32) hipLaunchKernelGGL(k_ysolve3<32>, g, b, 0, st, F);
    else if (mm <= 64) hipLaunchKernelGGL(k_ysolve3<64>, g, b, 0, st, F);
    else hipLaunchKernelGGL(k_ysolve3<128>, g, b, 0, st, F);
}

static int ysolve_maxm_b(const BatchSet& B)
{
    int mm = 0;
    for (const auto& F : B.hF) mm = std::max(mm, F.ks[0].maxm - 1);
    return mm;
}

void launch_ysolve1_b(BatchSet& B, hipStream_t st)
{
    const int mm = ysolve_maxm_b(B);
    const int c = y1c(mm <= 32 ? 32 : 64);
    const auto g1 = bdims(B, [c](const FluidDev& F) { return dim3((F.n2m + 63) / 64, (F.ks[0].nch + c - 1) / c); });
    const dim3 b1(64 * c);
    if (mm <= 16) blaunch(B, BK_Y1, g1, kb_ysolve1<16>, b1, 0, st);
    else if (mm <= 32) blaunch(B, BK_Y1, g1, kb_ysolve1<32>, b1, 0, st);
    else if (mm <= 64) blaunch(B, BK_Y1, g1, kb_ysolve1<64>, b1, 0, st);
    else blaunch(B, BK_Y1, g1, kb_ysolve1<128>, b1, 0, st);
}

void launch_ysolve_b(BatchSet& B, hipStream_t st, bool first_done)
{
    const int mm = ysolve_maxm_b(B);
    const dim3 b(64 * YW);
    const auto g = bdims(B, [](const FluidDev& F) { return dim3((F.n2m + 64 * YW - 1) / (64 * YW), F.ks[0].nch); });
    if (!first_done) launch_ysolve1_b(B, st);
    blaunch(B, BK_Y2, bdims(B, [](const FluidDev& F) { return dim3((F.n2m + YW - 1) / YW); }), kb_ysolve2, b, 0, st);
    if (mm <= 16) blaunch(B, BK_Y3, g, kb_ysolve3<16>, b, 0, st);
    else if (mm <= 32) blaunch(B, BK_Y3, g, kb_ysolve3<32>, b, 0, st);
    else if (mm <= 64) blaunch(B, BK_Y3, g, kb_ysolve3<64>, b, 0, st);
    else blaunch(B, BK_Y3, g, kb_ysolve3<128>, b, 0, st);
}
