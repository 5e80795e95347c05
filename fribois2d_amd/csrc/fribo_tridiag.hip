// Tridiagonal line solves of the fluid step (gfx950).
//
// j-direction (contiguous lines): solq2j / solq3j (solve_fluid.F90:4-103)
// and the per-mode Poisson solves of phcalc_outflow.F90:34-56.  One 64-lane
// wave owns one line, staged in LDS with coalesced loads/stores.  The line is
// split into up to 64 chunks (one per lane); each lane eliminates its chunk,
// the chunk separators form a 64-unknown tridiagonal system that the wave
// solves by parallel cyclic reduction over lane shuffles, and each lane then
// back-substitutes its chunk.  This is an exact direct solve of the same
// normalised system the reference hands to DGTTRF/DGTTRS (which never pivots
// on these diagonally dominant matrices); only the rounding order differs.
//
// k-direction (strided lines): solq2k / solq3k (solve_fluid.F90:107-255).
// Lanes run along j (coalesced), each thread owns one chunk of one k-line.
// The matrix does not depend on j, so every factor of the partitioned solve
// is computed once per substep (k_bc) and the per-line work is three passes:
// chunk elimination -> separator system (one wave per column, PCR with
// precomputed multipliers) -> chunk back-substitution fused with q += dq.
#include "fribo_internal.h"

size_t jsolve_lds_bytes(const FluidDev& F) { return (size_t)64 * F.jstride * sizeof(double); }

// Row r of a j-line belongs to chunk owner(r) when chunk starts are
// floor(p*n/nch).
__device__ inline int chunk_owner(int r, int n, int nch) { return (int)((((long)r + 1) * nch - 1) / n); }

// Coalesced global <-> LDS staging of one line in chunk layout: row t of
// chunk p at lds[p*stride + t] (odd stride: conflict-free lane access).
__device__ inline void line_to_lds(const double* __restrict__ src, double* lds, const int* cs, int n, int nch,
                                   int stride, int lane)
{
    for (int r = lane; r < n; r += 64) {
        const int p = chunk_owner(r, n, nch);
        lds[p * stride + (r - cs[p])] = src[r];
    }
}

__device__ inline void lds_to_line(double* __restrict__ dst, const double* lds, const int* cs, int n, int nch,
                                   int stride, int lane)
{
    for (int r = lane; r < n; r += 64) {
        const int p = chunk_owner(r, n, nch);
        dst[r] = lds[p * stride + (r - cs[p])];
    }
}

// Helmholtz j-lines (solq2j / solq3j): q2 for k = 1..n3m (rows j = 2..n2m;
// rows 1 and n2 are the Dirichlet identity rows with zero rhs), then q3 for
// k = 2..n3m (rows j = 1..n2m).  All factors are line-invariant (k_bc); the
// wave only runs the rhs through them.
template <int MAXR>
__global__ __launch_bounds__(64) void k_jsolve_h(const FluidDev F)
{
    extern __shared__ double lds[];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x, line = blockIdx.x;
    const int f = line < F.n3m ? 0 : 1;
    const int k = f == 0 ? line + 1 : line - F.n3m + 2;
    const KSys& K = F.js[f];
    const double* src = (f == 0 ? F.rhs2 : F.rhs3) + fidx(F, K.kofs, k);
    double* dst = (f == 0 ? F.sA : F.sB) + fidx(F, K.kofs, k);
    const int n = K.n, nch = K.nch, stride = F.jstride;
    line_to_lds(src, lds, K.cs, n, nch, stride, lane);
    __syncthreads();
    const int p = lane;
    double y[MAXR];
    double yl = 0.0, acc = 0.0, de = 0.0, ae = 0.0, ce = 0.0;
    int m = 0;
    double* row = lds + p * stride;
    if (p < nch) {
        m = K.cs[p + 1] - K.cs[p];
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            if (t < m - 1) {
                const int ix = t * 64 + p;
                yl = (row[t] * K.S[ix] - K.a[ix] * yl) * K.inv[ix];
                y[t] = yl;
                acc = acc + K.g[ix] * yl;
            }
        }
        const int ix = (m - 1) * 64 + p;
        de = row[m - 1] * K.S[ix];
        ae = K.a[ix];
        ce = K.c[ix];
    }
    const double ysn = __shfl_down(acc, 1, 64);
    double delta = (p < nch) ? de - ae * yl - ((p + 1 < nch) ? ce * ysn : 0.0) : 0.0;
    int li = 0;
    for (int h = 1; h < 64; h <<= 1, ++li) {
        double dm = __shfl_up(delta, h, 64), dp = __shfl_down(delta, h, 64);
        if (p - h < 0) dm = 0.0;
        if (p + h >= 64) dp = 0.0;
        delta = delta - K.k1[li * 64 + p] * dm - K.k2[li * 64 + p] * dp;
    }
    const double X = delta * K.ib[p];
    double Xm = __shfl_up(X, 1, 64);
    if (p == 0) Xm = 0.0;
    if (p < nch) {
        double x = 0.0;
#pragma unroll
        for (int t = MAXR - 1; t >= 0; --t) {
            if (t < m - 1) {
                const int ix = t * 64 + p;
                if (t == m - 2) x = y[t] + K.vp[ix] * Xm + K.wl[p] * X;
                else x = (y[t] + K.vp[ix] * Xm) - K.cp[ix] * x;
                row[t] = x;
            }
        }
        row[m - 1] = X;
    }
    __syncthreads();
    lds_to_line(dst, lds, K.cs, n, nch, stride, lane);
}

// Poisson j-lines (phcalc_outflow.F90:34-56): mode m solves
// (A_y + wr(m)) v = hmn(:,m); each lane factors its chunk in registers
// (unnormalised rows amphj, acphj+wr, apphj -- the same system the reference
// normalises by 1/(acphj+wr)), the separators are solved by PCR, and the
// singular Neumann mode is gauge-fixed by pinning its last unknown to 0.
template <int MAXR>
__global__ __launch_bounds__(64) void k_jsolve_p(const FluidDev F)
{
    extern __shared__ double lds[];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x, mode = blockIdx.x + 1;
    const PSys& Q = F.ps;
    double* col = F.hmn + (size_t)(mode - 1) * F.P;
    const double wrm = F.wr[mode];
    const bool pin = (mode == F.null_mode);
    const int n = Q.n, nch = Q.nch, stride = F.jstride;
    line_to_lds(col, lds, Q.cs, n, nch, stride, lane);
    __syncthreads();
    const int p = lane;
    double cp[MAXR], vp[MAXR], y[MAXR];
    double cpp = 0.0, yl = 0.0, vl = 1.0, gg = 1.0, ys = 0.0, vs = 0.0, ws = 0.0, wl = 0.0;
    double alpha = 0.0, beta = 1.0, gamma = 0.0, delta = 0.0, ae = 0.0, ce = 0.0, de = 0.0;
    int m = 0;
    double* row = lds + p * stride;
    if (p < nch) {
        m = Q.cs[p + 1] - Q.cs[p];
        double glast = 1.0;
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            if (t < m - 1) {
                const int ix = t * 64 + p;
                const double a = Q.am[ix], b = Q.ac[ix] + wrm, c = Q.ap[ix];
                const double iv = 1.0 / (b - a * cpp);
                cpp = c * iv;
                yl = (row[t] - a * yl) * iv;
                vl = (-a * vl) * iv;
                cp[t] = cpp;
                vp[t] = vl;
                y[t] = yl;
                ys = ys + gg * yl;
                vs = vs + gg * vl;
                if (t == m - 2) {
                    wl = (-c) * iv;
                    glast = gg;
                }
                gg = -cpp * gg;
            }
        }
        ws = glast * wl;
        const int ix = (m - 1) * 64 + p;
        ae = Q.am[ix];
        ce = Q.ap[ix];
        de = row[m - 1];
        alpha = ae * vl;
        beta = (Q.ac[ix] + wrm) + ae * wl;
        delta = de - ae * yl;
    }
    const double ysn = __shfl_down(ys, 1, 64), vsn = __shfl_down(vs, 1, 64), wsn = __shfl_down(ws, 1, 64);
    if (p + 1 < nch) {
        beta = beta + ce * vsn;
        gamma = ce * wsn;
        delta = delta - ce * ysn;
    }
    if (pin && p == nch - 1) {
        alpha = 0.0;
        beta = 1.0;
        gamma = 0.0;
        delta = 0.0;
    }
    for (int h = 1; h < 64; h <<= 1) {
        double am = __shfl_up(alpha, h, 64), bm = __shfl_up(beta, h, 64);
        double gm = __shfl_up(gamma, h, 64), dm = __shfl_up(delta, h, 64);
        double ap = __shfl_down(alpha, h, 64), bp = __shfl_down(beta, h, 64);
        double gp = __shfl_down(gamma, h, 64), dp = __shfl_down(delta, h, 64);
        if (lane - h < 0) { am = 0.0; bm = 1.0; gm = 0.0; dm = 0.0; }
        if (lane + h >= 64) { ap = 0.0; bp = 1.0; gp = 0.0; dp = 0.0; }
        const double m1 = alpha / bm, m2 = gamma / bp;
        const double na = -m1 * am, ng = -m2 * gp;
        const double nb = beta - m1 * gm - m2 * ap;
        const double nd = delta - m1 * dm - m2 * dp;
        alpha = na;
        beta = nb;
        gamma = ng;
        delta = nd;
    }
    const double X = delta / beta;
    double Xm = __shfl_up(X, 1, 64);
    if (p == 0) Xm = 0.0;
    if (p < nch) {
        double x = 0.0;
#pragma unroll
        for (int t = MAXR - 1; t >= 0; --t) {
            if (t < m - 1) {
                if (t == m - 2) x = y[t] + vp[t] * Xm + wl * X;
                else x = (y[t] + vp[t] * Xm) - cp[t] * x;
                row[t] = x;
            }
        }
        row[m - 1] = X;
    }
    __syncthreads();
    lds_to_line(col, lds, Q.cs, n, nch, stride, lane);
}

void launch_jsolve_helm(const FluidDev& F, hipStream_t st)
{
    const int nlines = F.n3m + (F.n3m - 1);
    const int mr = (F.js[0].maxm > F.js[1].maxm ? F.js[0].maxm : F.js[1].maxm) - 1;
    const size_t lds = jsolve_lds_bytes(F);
    if (mr <= 8) hipLaunchKernelGGL(k_jsolve_h<8>, dim3(nlines), dim3(64), lds, st, F);
    else if (mr <= 16) hipLaunchKernelGGL(k_jsolve_h<16>, dim3(nlines), dim3(64), lds, st, F);
    else if (mr <= 24) hipLaunchKernelGGL(k_jsolve_h<24>, dim3(nlines), dim3(64), lds, st, F);
    else if (mr <= 32) hipLaunchKernelGGL(k_jsolve_h<32>, dim3(nlines), dim3(64), lds, st, F);
    else hipLaunchKernelGGL(k_jsolve_h<64>, dim3(nlines), dim3(64), lds, st, F);
}

void launch_jsolve_poisson(const FluidDev& F, hipStream_t st)
{
    const int mr = F.ps.maxm - 1;
    const size_t lds = jsolve_lds_bytes(F);
    if (mr <= 8) hipLaunchKernelGGL(k_jsolve_p<8>, dim3(F.n3m), dim3(64), lds, st, F);
    else if (mr <= 16) hipLaunchKernelGGL(k_jsolve_p<16>, dim3(F.n3m), dim3(64), lds, st, F);
    else if (mr <= 24) hipLaunchKernelGGL(k_jsolve_p<24>, dim3(F.n3m), dim3(64), lds, st, F);
    else if (mr <= 32) hipLaunchKernelGGL(k_jsolve_p<32>, dim3(F.n3m), dim3(64), lds, st, F);
    else hipLaunchKernelGGL(k_jsolve_p<64>, dim3(F.n3m), dim3(64), lds, st, F);
}

// ---------------------------------------------------------------------------
// k-direction solves.  Field f = 0: q2 (rows kc=1..n3m), f = 1: q3 (interior
// rows kc=2..n3m of the n3 system; the Dirichlet rows dqb3s/dqb3n enter the
// first/last rhs).
__device__ inline double krhs(const FluidDev& F, const KSys& K, int f, int j, int r, const double* src,
                              const DevScalars* S)
{
    double d = src[fidx(F, j, r + K.kofs)] * K.S[r];
    if (f == 0) {
        if (r == 0) d = d + F.dqb2s[j] * S->ugkks;
        if (r == K.n - 1) d = d + F.dqb2n[j] * S->ugkkn;
    } else {
        if (r == 0) d = d - K.bc[0] * F.dqb3s[j];
        if (r == K.n - 1) d = d - K.bc[1] * F.dqb3n[j];
    }
    return d;
}

__global__ __launch_bounds__(64) void k_ksolve1(const FluidDev F)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = blockIdx.z, p = blockIdx.y;
    const KSys& K = F.ks[f];
    const int j = blockIdx.x * 64 + threadIdx.x + 1;
    if (p >= K.nch || j > F.n2m) return;
    const double* src = f == 0 ? F.sA : F.sB;
    const int s = K.cs[p], e = K.cs[p + 1] - 1;
    double yp = 0.0, acc = 0.0;
    for (int i = s; i <= e - 1; ++i) {
        double d = krhs(F, K, f, j, i, src, S);
        yp = (d - K.a[i] * yp) * K.inv[i];
        acc = acc + K.g[i] * yp;
    }
    double de = krhs(F, K, f, j, e, src, S);
    const size_t o = ((size_t)f * FRIBO_KMAXCH + p) * F.n2m + (j - 1);
    F.kYs[o] = acc;
    F.kYe[o] = yp;
    F.kDe[o] = de;
}

__global__ __launch_bounds__(64) void k_ksolve2(const FluidDev F)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = blockIdx.y, j = blockIdx.x + 1, p = threadIdx.x;
    const KSys& K = F.ks[f];
    const size_t base = (size_t)f * FRIBO_KMAXCH * F.n2m + (j - 1);
    double delta = 0.0;
    if (p < K.nch) {
        int e = K.cs[p + 1] - 1;
        double ysn = (p + 1 < K.nch) ? F.kYs[base + (size_t)(p + 1) * F.n2m] : 0.0;
        double ye = F.kYe[base + (size_t)p * F.n2m];
        double de = F.kDe[base + (size_t)p * F.n2m];
        delta = de - K.a[e] * ye - K.c[e] * ysn;
    }
    int li = 0;
    for (int h = 1; h < 64; h <<= 1, ++li) {
        double dm = __shfl_up(delta, h, 64), dp = __shfl_down(delta, h, 64);
        if (p - h < 0) dm = 0.0;
        if (p + h >= 64) dp = 0.0;
        delta = delta - K.k1[li * 64 + p] * dm - K.k2[li * 64 + p] * dp;
    }
    if (p < K.nch) F.kX[base + (size_t)p * F.n2m] = delta * K.ib[p];
}

template <int MAXM>
__global__ __launch_bounds__(64) void k_ksolve3(const FluidDev F)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = blockIdx.z, p = blockIdx.y;
    const KSys& K = F.ks[f];
    const int j = blockIdx.x * 64 + threadIdx.x + 1;
    if (p >= K.nch || j > F.n2m) return;
    const double* __restrict__ src = f == 0 ? F.sA : F.sB;
    double* __restrict__ q = f == 0 ? F.q2 : F.q3;
    const int s = K.cs[p], e = K.cs[p + 1] - 1, m = e - s;   // m interior rows
    const size_t base = (size_t)f * FRIBO_KMAXCH * F.n2m + (j - 1);
    const double X = F.kX[base + (size_t)p * F.n2m];
    const double Xm = (p > 0) ? F.kX[base + (size_t)(p - 1) * F.n2m] : 0.0;
    const bool zero_row = (f == 0 && j == 1);
    double y[MAXM], qv[MAXM];
    double yp = 0.0;
#pragma unroll
    for (int t = 0; t < MAXM; ++t) {
        if (t < m) {
            const int i = s + t;
            const double d = krhs(F, K, f, j, i, src, S);
            qv[t] = q[fidx(F, j, i + K.kofs)];
            yp = (d - K.a[i] * yp) * K.inv[i];
            y[t] = yp;
        }
    }
    const size_t ce = fidx(F, j, e + K.kofs);
    const double qe = q[ce];
    double x = 0.0;
#pragma unroll
    for (int t = MAXM - 1; t >= 0; --t) {
        if (t < m) {
            const int i = s + t;
            if (t == m - 1) x = y[t] + K.vp[i] * Xm + K.wl[p] * X;
            else x = (y[t] + K.vp[i] * Xm) - K.cp[i] * x;
            q[fidx(F, j, i + K.kofs)] = zero_row ? 0.0 : qv[t] + x;
            if (zero_row) q[fidx(F, F.n2, i + K.kofs)] = 0.0;
        }
    }
    q[ce] = zero_row ? 0.0 : qe + X;
    if (zero_row) q[fidx(F, F.n2, e + K.kofs)] = 0.0;
    if (f == 1) {
        if (p == 0) {
            const size_t c = fidx(F, j, 1);
            q[c] = q[c] + F.dqb3s[j];
        }
        if (p == K.nch - 1) {
            const size_t c = fidx(F, j, F.n3);
            q[c] = q[c] + F.dqb3n[j];
        }
    }
}

void launch_ksolve(const FluidDev& F, hipStream_t st)
{
    int nchmax = F.ks[0].nch > F.ks[1].nch ? F.ks[0].nch : F.ks[1].nch;
    int maxm = F.ks[0].maxm > F.ks[1].maxm ? F.ks[0].maxm : F.ks[1].maxm;
    dim3 g((F.n2m + 63) / 64, nchmax, 2);
    hipLaunchKernelGGL(k_ksolve1, g, dim3(64), 0, st, F);
    hipLaunchKernelGGL(k_ksolve2, dim3(F.n2m, 2), dim3(64), 0, st, F);
    if (maxm <= 8) hipLaunchKernelGGL(k_ksolve3<8>, g, dim3(64), 0, st, F);
    else if (maxm <= 16) hipLaunchKernelGGL(k_ksolve3<16>, g, dim3(64), 0, st, F);
    else if (maxm <= 32) hipLaunchKernelGGL(k_ksolve3<32>, g, dim3(64), 0, st, F);
    else if (maxm <= 64) hipLaunchKernelGGL(k_ksolve3<64>, g, dim3(64), 0, st, F);
    else hipLaunchKernelGGL(k_ksolve3<128>, g, dim3(64), 0, st, F);
}
