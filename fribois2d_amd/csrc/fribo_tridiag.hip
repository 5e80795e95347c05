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

size_t jsolve_lds_bytes(int n) { return (size_t)3 * n * sizeof(double); }

// Coefficients of row r of a j-line (normalised diagonal 1).
struct JCoef {
    const double *am, *ac, *ap;
    double s0, s1, s2;   // S = 1/(s0 + s1*ac[j]),  a = (s2*am[j])*S,  c = (s2*ap[j])*S
    int jofs;            // j index of row 0
    int n;
    __device__ inline void row(int r, double& a, double& c, double& S) const
    {
        int j = r + jofs;
        S = 1.0 / (s0 + s1 * ac[j]);
        a = (r == 0) ? 0.0 : s2 * am[j] * S;
        c = (r == n - 1) ? 0.0 : s2 * ap[j] * S;
    }
};

// Partitioned solve of one line held in LDS (D: scaled rhs in, solution out).
__device__ void wave_line_solve(const JCoef& Q, double* C, double* D, double* V, int lane, bool pin_last)
{
    const int n = Q.n;
    int P = n / 2;
    if (P > 64) P = 64;
    if (P < 1) P = 1;
    const int p = lane;
    const bool act = p < P;
    double yp = 0.0, vpp = 1.0, wl = 0.0, ys = 0.0, vs = 0.0, ws = 0.0;
    double alpha = 0.0, beta = 1.0, gamma = 0.0, delta = 0.0;
    double ae = 0.0, ce = 0.0, de = 0.0;
    int s = 0, e = 0;
    if (act) {
        s = (int)(((long)p * n) / P);
        e = (int)(((long)(p + 1) * n) / P) - 1;
        double cpp = 0.0;
        vpp = 1.0;
        yp = 0.0;
        for (int i = s; i <= e - 1; ++i) {
            double a, c, S;
            Q.row(i, a, c, S);
            double iv = 1.0 / (1.0 - a * cpp);
            double cpi = c * iv;
            double ypi = (D[i] - a * yp) * iv;
            double vpi = (-a * vpp) * iv;
            C[i] = cpi;
            D[i] = ypi;
            V[i] = vpi;
            cpp = cpi;
            yp = ypi;
            vpp = vpi;
            if (i == e - 1) wl = (-c) * iv;
        }
        double yb = yp, vb = vpp, wb = wl;
        for (int i = e - 2; i >= s; --i) {
            yb = D[i] - C[i] * yb;
            vb = V[i] - C[i] * vb;
            wb = -C[i] * wb;
        }
        ys = yb;
        vs = vb;
        ws = wb;
        double S;
        Q.row(e, ae, ce, S);
        de = D[e];
    }
    double ys_n = __shfl_down(ys, 1, 64);
    double vs_n = __shfl_down(vs, 1, 64);
    double ws_n = __shfl_down(ws, 1, 64);
    if (act) {
        const bool nxt = (p + 1) < P;
        alpha = ae * vpp;
        beta = 1.0 + ae * wl + (nxt ? ce * vs_n : 0.0);
        gamma = nxt ? ce * ws_n : 0.0;
        delta = de - ae * yp - (nxt ? ce * ys_n : 0.0);
        if (pin_last && p == P - 1) {
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
        double m1 = alpha / bm, m2 = gamma / bp;
        double na = -m1 * am, ng = -m2 * gp;
        double nb = beta - m1 * gm - m2 * ap;
        double nd = delta - m1 * dm - m2 * dp;
        alpha = na;
        beta = nb;
        gamma = ng;
        delta = nd;
    }
    double X = delta / beta;
    double Xm = __shfl_up(X, 1, 64);
    if (p == 0) Xm = 0.0;
    if (act) {
        double x = yp + vpp * Xm + wl * X;
        D[e - 1] = x;
        for (int i = e - 2; i >= s; --i) {
            x = (D[i] + V[i] * Xm) - C[i] * x;
            D[i] = x;
        }
        D[e] = X;
    }
}

// MODE 0: Helmholtz lines (q2 for k=1..n3m, then q3 for k=2..n3m).
// MODE 1: Poisson modes m=1..n3m on the transformed field hmn.
template <int MODE>
__global__ __launch_bounds__(64) void k_jsolve(const FluidDev F)
{
    extern __shared__ double lds[];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x;
    const int line = blockIdx.x;
    JCoef Q;
    const double* src;
    double* dst;
    bool pin = false;
    if (MODE == 0) {
        const double betadx = S->betadx;
        Q.s0 = 1.0;
        Q.s1 = -betadx;
        Q.s2 = -betadx;
        if (line < F.n3m) {
            int k = line + 1;
            Q.am = F.am2j; Q.ac = F.ac2j; Q.ap = F.ap2j;
            Q.jofs = 2;
            Q.n = F.n2 - 2;
            src = F.rhs2 + fidx(F, 2, k);
            dst = F.sA + fidx(F, 2, k);
        } else {
            int k = line - F.n3m + 2;
            Q.am = F.am3j; Q.ac = F.ac3j; Q.ap = F.ap3j;
            Q.jofs = 1;
            Q.n = F.n2m;
            src = F.rhs3 + fidx(F, 1, k);
            dst = F.sB + fidx(F, 1, k);
        }
    } else {
        const int m = line + 1;
        Q.s0 = F.wr[m];
        Q.s1 = 1.0;
        Q.s2 = 1.0;
        Q.am = F.amphj; Q.ac = F.acphj; Q.ap = F.apphj;
        Q.jofs = 1;
        Q.n = F.n2m;
        src = F.hmn + (size_t)line * F.P;
        dst = F.hmn + (size_t)line * F.P;
        pin = (m == F.null_mode);
    }
    const int n = Q.n;
    double* C = lds;
    double* D = lds + n;
    double* V = lds + 2 * n;
    for (int r = lane; r < n; r += 64) {
        double a, c, Sr;
        Q.row(r, a, c, Sr);
        D[r] = src[r] * Sr;
    }
    __syncthreads();
    wave_line_solve(Q, C, D, V, lane, pin);
    __syncthreads();
    for (int r = lane; r < n; r += 64) dst[r] = D[r];
}

void launch_jsolve_helm(const FluidDev& F, hipStream_t st)
{
    int nlines = F.n3m + (F.n3m - 1);
    hipLaunchKernelGGL(k_jsolve<0>, dim3(nlines), dim3(64), jsolve_lds_bytes(F.n2), st, F);
}

void launch_jsolve_poisson(const FluidDev& F, hipStream_t st)
{
    hipLaunchKernelGGL(k_jsolve<1>, dim3(F.n3m), dim3(64), jsolve_lds_bytes(F.n2m), st, F);
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
    const double* src = f == 0 ? F.sA : F.sB;
    double* q = f == 0 ? F.q2 : F.q3;
    const int s = K.cs[p], e = K.cs[p + 1] - 1, m = e - s;   // m interior rows
    const size_t base = (size_t)f * FRIBO_KMAXCH * F.n2m + (j - 1);
    const double X = F.kX[base + (size_t)p * F.n2m];
    const double Xm = (p > 0) ? F.kX[base + (size_t)(p - 1) * F.n2m] : 0.0;
    double y[MAXM];
    double yp = 0.0;
#pragma unroll
    for (int t = 0; t < MAXM; ++t) {
        if (t < m) {
            int i = s + t;
            double d = krhs(F, K, f, j, i, src, S);
            yp = (d - K.a[i] * yp) * K.inv[i];
            y[t] = yp;
        }
    }
    const bool zero_row = (f == 0 && j == 1);
    double x = 0.0;
#pragma unroll
    for (int t = MAXM - 1; t >= 0; --t) {
        if (t < m) {
            int i = s + t;
            if (t == m - 1) x = y[t] + K.vp[i] * Xm + K.wl[p] * X;
            else x = (y[t] + K.vp[i] * Xm) - K.cp[i] * x;
            size_t c = fidx(F, j, i + K.kofs);
            q[c] = zero_row ? 0.0 : q[c] + x;
            if (f == 0 && j == 1) q[fidx(F, F.n2, i + K.kofs)] = 0.0;
        }
    }
    {
        size_t c = fidx(F, j, e + K.kofs);
        q[c] = zero_row ? 0.0 : q[c] + X;
        if (f == 0 && j == 1) q[fidx(F, F.n2, e + K.kofs)] = 0.0;
    }
    if (f == 1) {
        if (p == 0) {
            size_t c = fidx(F, j, 1);
            q[c] = q[c] + F.dqb3s[j];
        }
        if (p == K.nch - 1) {
            size_t c = fidx(F, j, F.n3);
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
