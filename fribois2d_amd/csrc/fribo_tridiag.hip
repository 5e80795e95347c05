// Tridiagonal line solves of the fluid step (gfx950).
//
// j-direction (contiguous lines): solq2j / solq3j (solve_fluid.F90:4-103)
// and the per-mode Poisson solves of phcalc_outflow.F90:34-56.  A workgroup
// of JW waves keeps the line-invariant data of one system in LDS (the
// Helmholtz factors, or the Poisson coefficient rows) and its waves loop over
// lines, one line per wave at a time.  A line is split into up to 64 chunks
// (one per lane): the wave loads the line coalesced into registers (the next
// line's loads are in flight while the current line is solved), scatters it
// into an LDS chunk layout, each lane eliminates its chunk, the chunk
// separators form a 64-unknown tridiagonal system that the wave solves by
// parallel cyclic reduction over lane shuffles, each lane back-substitutes its
// chunk, and the line is gathered back coalesced.  This is an exact direct
// solve of the same normalised system the reference hands to DGTTRF/DGTTRS
// (which never pivots on these diagonally dominant matrices); only the
// rounding order differs.
//
// A lane's m-1 interior rows are aligned to the end of a MAXR-step loop
// (front padding), so every step reads its LDS data unconditionally (clamped
// index); padded steps leave the recurrences untouched.
//
// k-direction (strided lines): solq2k / solq3k (solve_fluid.F90:107-255).
// Lanes run along j (coalesced), each thread owns one chunk of one k-line.
// The matrix does not depend on j, so every factor of the partitioned solve
// is computed once per substep (k_bc) and the per-line work is three passes:
// chunk elimination -> separator system (one wave per column, PCR with
// precomputed multipliers) -> chunk back-substitution fused with q += dq.
#include "fribo_batch.h"

#include <algorithm>
#include <utility>

// j-lines per wave of the batched Helmholtz j-line launch
#ifndef FRIBO_JB_LPW
#define FRIBO_JB_LPW 16
#endif

// Row r of a j-line belongs to chunk owner(r) when chunk starts are
// floor(p*n/nch); (r+1)*nch < 2^32 for every line this build accepts.
__device__ inline int chunk_owner(int r, int n, int nch)
{
    return (int)(((unsigned)(r + 1) * (unsigned)nch - 1u) / (unsigned)n);
}

// Per-lane staging map of a j-line: element u of the lane (row r = lane+64u)
// goes to LDS slot p*stride + (r - cs[p]) of the chunk layout (-1: no row).
// Chunk starts come from the lanes (lane p holds cs[p]) by shuffles.
template <int NR>
__device__ inline void line_map(int* slot, const int* cs, int n, int nch, int stride, int lane)
{
    const int csl = lane < nch ? cs[lane] : 0;
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int r = lane + 64 * u;
        const int p = chunk_owner(r < n ? r : n - 1, n, nch);
        const int c0 = __shfl(csl, p, 64);
        slot[u] = r < n ? p * stride + (r - c0) : -1;
    }
}

// Staging map with every chunk front-padded to a common length: the m-1
// interior rows of chunk p sit at offsets off_p .. MAXR-1 of its staging row
// (off_p = MAXR - (m_p - 1)) and the separator at offset MAXR, so a solve step
// tt reads offset tt of every lane with no per-lane clamping.
template <int NR>
__device__ inline void line_map_pad(int* slot, const int* cs, int n, int nch, int stride, int lane, int maxr)
{
    const int csl = lane < nch ? cs[lane] : 0;
    const int csn = lane < nch ? cs[lane + 1] : 0;
    const int offl = maxr - (csn - csl - 1);
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int r = lane + 64 * u;
        const int p = chunk_owner(r < n ? r : n - 1, n, nch);
        const int c0 = __shfl(csl, p, 64), o = __shfl(offl, p, 64);
        slot[u] = r < n ? p * stride + (r - c0) + o : -1;
    }
}

// Loads of a line's rows (lane + 64u), clamped to the last row so every load
// is unconditional; rows past the line are never staged, and the caller
// passes a valid line for an inactive wave.
template <int NR>
__device__ inline void line_fetch(double* pre, const double* __restrict__ src, int n, int lane)
{
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int r = lane + 64 * u;
        pre[u] = src[r < n ? r : n - 1];
    }
}

template <int NR>
__device__ inline void line_put(double* stage, const double* pre, const int* slot)
{
#pragma unroll
    for (int u = 0; u < NR; ++u)
        if (slot[u] >= 0) stage[slot[u]] = pre[u];
}

template <int NR>
__device__ inline void line_get(double* __restrict__ dst, const double* stage, const int* slot, int lane)
{
#pragma unroll
    for (int u = 0; u < NR; ++u)
        if (slot[u] >= 0) dst[lane + 64 * u] = stage[slot[u]];
}

// The staging rows of a j-line are private to its wave, so the line loop
// only has to order the wave's own LDS accesses across its lanes (LDS
// executes a wave's accesses in order).  A workgroup barrier here would also
// wait for the next line's global loads, serialising load and solve.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

size_t jsolve_h_lds_bytes(int maxr, int jw, int stride, bool fl)
{
    const size_t nr = (size_t)maxr + 1;
    return ((fl ? 4 * nr * 64 + 2 * 384 + 2 * 64 : 0) + (size_t)jw * 64 * stride) * sizeof(double);
}

size_t jsolve_p_lds_bytes(int maxr, int jw, int stride)
{
    const size_t nr = (size_t)maxr + 1;
    return (3 * nr * 64 + (size_t)jw * 64 * stride) * sizeof(double);
}

// Helmholtz j-lines (solq2j / solq3j): q2 for k = 1..n3m (rows j = 2..n2m;
// rows 1 and n2 are the Dirichlet identity rows with zero rhs), then q3 for
// k = 2..n3m (rows j = 1..n2m).  All factors are line-invariant (k_bc); the
// first half of the grid takes the q2 lines, the second half the q3 lines.
//
// The workgroup copies the factors into LDS front-padded per chunk like the
// staged line, as four arrays: S*inv and a*inv (the forward step becomes
// y' = d*(S*inv) - (a*inv)*y'; the separator row keeps S and a themselves),
// vp and cp.  The weights g of the separator sums are regenerated on the fly
// (g_0 = 1, g_{t+1} = -cp_t g_t, the recurrence that produced them).  Padding
// rows have S*inv = a*inv = vp = 0 and cp = -1, which keeps y' at zero and g
// at one until the chunk's first row, so every step tt reads row tt of every
// lane without clamps or selects.  Four arrays leave room for eight waves
// (two per SIMD) on grids up to n2 = 1537.
template <int MAXR, int JW>
__device__ __forceinline__ void jsolve_hp_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int NR = MAXR + 1;
    extern __shared__ double lds[];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = (int)vb.gx >> 1;
    const int f = (int)vb.x < half ? 0 : 1;
    const int wg = f == 0 ? (int)vb.x : (int)vb.x - half;
    const int nwgf = f == 0 ? half : (int)vb.gx - half;
    const KSys& K = F.js[f];
    const int nlines = f == 0 ? F.n3m : F.n3m - 1;
    const int kfirst = f == 0 ? 1 : 2;
    const double* srcb = f == 0 ? F.rhs2 : F.rhs3;
    double* dstb = f == 0 ? F.sA : F.sB;
    const int n = K.n, nch = K.nch, stride = F.jstride;
    const int p = lane;
    const bool own = p < nch;
    const int m = own ? K.cs[p + 1] - K.cs[p] : 1;
    const int ie = (m - 1) * 64 + p;

    double* fSI = lds;
    double* fAI = fSI + NR * 64;
    double* fv = fAI + NR * 64;
    double* fc = fv + NR * 64;
    double* fk1 = fc + NR * 64;
    double* fk2 = fk1 + 384;
    double* fwl = fk2 + 384;
    double* fib = fwl + 64;
    double* stage = fib + 64 + wave * 64 * stride;
    {
        // padded copy: LDS row tt of lane q holds factor row tt - off_q of
        // chunk q (padding where that is negative); all loads before the stores
        constexpr int NCP = (NR * 64 + 64 * JW - 1) / (64 * JW);
        constexpr int NCK = (384 + 64 * JW - 1) / (64 * JW);
        double tS[NCP], ta[NCP], ti[NCP], tv[NCP], tc[NCP], t1[NCK], t2[NCK];
        bool ok[NCP];
#pragma unroll
        for (int u = 0; u < NCP; ++u) {
            const int i = tid + u * 64 * JW;
            const int q = i & 63, tt = i >> 6;
            const int mq = q < nch ? K.cs[q + 1] - K.cs[q] : 0;
            const int t = tt - MAXR + mq - 1;
            ok[u] = i < NR * 64 && mq > 0 && t >= 0;
            const int ic = ok[u] ? t * 64 + q : 0;
            tS[u] = K.S[ic];
            ta[u] = K.a[ic];
            ti[u] = K.inv[ic];
            tv[u] = K.vp[ic];
            tc[u] = K.cp[ic];
        }
#pragma unroll
        for (int u = 0; u < NCK; ++u) {
            const int i = tid + u * 64 * JW;
            t1[u] = K.k1[i < 384 ? i : 0];
            t2[u] = K.k2[i < 384 ? i : 0];
        }
        const double twl = K.wl[lane], tib = K.ib[lane];
#pragma unroll
        for (int u = 0; u < NCP; ++u) {
            const int i = tid + u * 64 * JW;
            if (i < NR * 64) {
                const bool sep = (i >> 6) == MAXR;   // separator row: S and a themselves
                fSI[i] = ok[u] ? (sep ? tS[u] : tS[u] * ti[u]) : 0.0;
                fAI[i] = ok[u] ? (sep ? ta[u] : ta[u] * ti[u]) : 0.0;
                fv[i] = ok[u] && !sep ? tv[u] : 0.0;
                fc[i] = ok[u] ? (sep ? 0.0 : tc[u]) : -1.0;
            }
        }
#pragma unroll
        for (int u = 0; u < NCK; ++u) {
            const int i = tid + u * 64 * JW;
            if (i < 384) {
                fk1[i] = t1[u];
                fk2[i] = t2[u];
            }
        }
        if (tid < 64) {
            fwl[tid] = twl;
            fib[tid] = tib;
        }
        // the padding rows of the staging area start at zero
        for (int i = lane; i < 64 * stride; i += 64) stage[i] = 0.0;
    }
    const double ce = own ? K.c[ie] : 0.0;
    int slot[NR];
    line_map_pad<NR>(slot, K.cs, n, nch, stride, lane, MAXR);
    const int step = nwgf * JW;
    const int iters = (nlines + step - 1) / step;
    int line = wg * JW + wave;
    double* row = stage + p * stride;
    // one line: stage it, refill its buffer with the line `ahead` lines of
    // this wave later (one line of loads in flight: a second buffer takes
    // the kernel past 256 VGPRs), solve, write back
    auto solve = [&](double (&pre)[NR], int ahead) {
        const bool act = line < nlines;
        line_put<NR>(stage, pre, slot);
        const int nxt = line + ahead * step;
        line_fetch<NR>(pre, srcb + fidx(F, K.kofs, kfirst + (nxt < nlines ? nxt : 0)), n, lane);
        wave_lds_sync();
        double y[MAXR];
        double yl = 0.0, acc = 0.0, gg = 1.0;
#pragma unroll
        for (int tt = 0; tt < MAXR; ++tt) {
            const int ix = tt * 64 + p;
            yl = row[tt] * fSI[ix] - fAI[ix] * yl;
            y[tt] = yl;
            acc = acc + gg * yl;
            gg = -fc[ix] * gg;
        }
        const double de = row[MAXR] * fSI[MAXR * 64 + p];
        const double ae = fAI[MAXR * 64 + p];
        const double ysn = __shfl_down(acc, 1, 64);
        double delta = own ? de - ae * yl - ((p + 1 < nch) ? ce * ysn : 0.0) : 0.0;
        int li = 0;
        for (int h = 1; h < 64; h <<= 1, ++li) {
            double dm = __shfl_up(delta, h, 64), dp = __shfl_down(delta, h, 64);
            if (p - h < 0) dm = 0.0;
            if (p + h >= 64) dp = 0.0;
            delta = delta - fk1[li * 64 + p] * dm - fk2[li * 64 + p] * dp;
        }
        const double X = delta * fib[p];
        double Xm = __shfl_up(X, 1, 64);
        if (p == 0) Xm = 0.0;
        const double wl = fwl[p];
        double x = 0.0;
#pragma unroll
        for (int tt = MAXR - 1; tt >= 0; --tt) {
            const int ix = tt * 64 + p;
            if (tt == MAXR - 1) x = y[tt] + fv[ix] * Xm + wl * X;
            else x = (y[tt] + fv[ix] * Xm) - fc[ix] * x;
            row[tt] = x;
        }
        row[MAXR] = X;
        wave_lds_sync();
        if (act) line_get<NR>(dstb + fidx(F, K.kofs, kfirst + line), stage, slot, lane);
        wave_lds_sync();
        line += step;
    };
    double pre[NR];
    line_fetch<NR>(pre, srcb + fidx(F, K.kofs, kfirst + (line < nlines ? line : 0)), n, lane);
    __syncthreads();
    for (int it = 0; it < iters; ++it) solve(pre, 1);
}
template <int MAXR, int JW>
__global__ __launch_bounds__(64 * JW) void k_jsolve_hp(const FluidDev F)
{
    jsolve_hp_body<MAXR, JW>(F, vb_hw());
}
template <int MAXR, int JW>
__global__ __launch_bounds__(64 * JW) void kb_jsolve_hp(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    jsolve_hp_body<MAXR, JW>(*Fp, vb);
}

// The same solve with the factors read from global memory (grids whose chunk
// factors do not fit LDS).
template <int MAXR>
__device__ __forceinline__ void jsolve_h_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int NR = MAXR + 1;
    extern __shared__ double lds[];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x & 63;
    const int half = (int)vb.gx >> 1;
    const int f = (int)vb.x < half ? 0 : 1;
    const int wg = f == 0 ? (int)vb.x : (int)vb.x - half;
    const int nwgf = f == 0 ? half : (int)vb.gx - half;
    const KSys& K = F.js[f];
    const int nlines = f == 0 ? F.n3m : F.n3m - 1;
    const int kfirst = f == 0 ? 1 : 2;
    const double* srcb = f == 0 ? F.rhs2 : F.rhs3;
    double* dstb = f == 0 ? F.sA : F.sB;
    const int n = K.n, nch = K.nch, stride = F.jstride;
    const int p = lane;
    const bool own = p < nch;
    const int m = own ? K.cs[p + 1] - K.cs[p] : 1;
    const int off = MAXR - (m - 1);
    const int ie = (m - 1) * 64 + p;
    const double *fS = K.S, *fa = K.a, *fi = K.inv, *fg = K.g, *fv = K.vp, *fc = K.cp;
    double* stage = lds;
    const double ce = own ? K.c[ie] : 0.0;
    int slot[NR];
    line_map<NR>(slot, K.cs, n, nch, stride, lane);
    double pre[NR];
    const int iters = (nlines + nwgf - 1) / nwgf;
    int line = wg;
    line_fetch<NR>(pre, srcb + fidx(F, K.kofs, kfirst + (line < nlines ? line : 0)), n, lane);
    double* row = stage + p * stride;
    for (int it = 0; it < iters; ++it, line += nwgf) {
        const bool act = line < nlines;
        line_put<NR>(stage, pre, slot);
        const int nxt = line + nwgf;
        line_fetch<NR>(pre, srcb + fidx(F, K.kofs, kfirst + (nxt < nlines ? nxt : 0)), n, lane);
        wave_lds_sync();
        double y[MAXR];
        double yl = 0.0, acc = 0.0;
#pragma unroll
        for (int tt = 0; tt < MAXR; ++tt) {
            const int t = tt - off;
            const bool real = t >= 0;
            const int tc = real ? t : 0;
            const int ix = tc * 64 + p;
            const double sv = real ? fS[ix] : 0.0;
            const double av = real ? fa[ix] : 0.0;
            const double gv = real ? fg[ix] : 0.0;
            const double dv = real ? row[tc] : 0.0;
            yl = (dv * sv - av * yl) * fi[ix];
            y[tt] = yl;
            acc = acc + gv * yl;
        }
        const double de = row[m - 1] * fS[ie];
        const double ae = fa[ie];
        const double ysn = __shfl_down(acc, 1, 64);
        double delta = own ? de - ae * yl - ((p + 1 < nch) ? ce * ysn : 0.0) : 0.0;
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
        const double wl = K.wl[p];
        double x = 0.0;
#pragma unroll
        for (int tt = MAXR - 1; tt >= 0; --tt) {
            const int t = tt - off;
            const int ix = (t >= 0 ? t : 0) * 64 + p;
            if (tt == MAXR - 1) x = y[tt] + fv[ix] * Xm + wl * X;
            else x = (y[tt] + fv[ix] * Xm) - fc[ix] * x;
            if (t >= 0 && own) row[t] = x;
        }
        if (own) row[m - 1] = X;
        wave_lds_sync();
        if (act) line_get<NR>(dstb + fidx(F, K.kofs, kfirst + line), stage, slot, lane);
        wave_lds_sync();
    }
}
template <int MAXR>
__global__ __launch_bounds__(64) void k_jsolve_h(const FluidDev F)
{
    jsolve_h_body<MAXR>(F, vb_hw());
}
template <int MAXR>
__global__ __launch_bounds__(64) void kb_jsolve_h(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    jsolve_h_body<MAXR>(*Fp, vb);
}

// Poisson j-lines (phcalc_outflow.F90:34-56): mode m solves
// (A_y + wr(m)) v = hmn(:,m); each lane factors its chunk in registers
// (unnormalised rows amphj, acphj+wr, apphj from LDS -- the same system the
// reference normalises by 1/(acphj+wr)), the separators are solved by PCR,
// and the singular Neumann mode is gauge-fixed by pinning its last unknown
// to 0.
template <int MAXR, int JW>
__device__ __forceinline__ void jsolve_p_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int NR = MAXR + 1;
    extern __shared__ double lds[];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const PSys& Q = F.ps;
    const int n = Q.n, nch = Q.nch, stride = F.jstride;
    double* cam = lds;
    double* cac = cam + NR * 64;
    double* cap = cac + NR * 64;
    double* stage = cap + NR * 64 + wave * 64 * stride;
    const int lenf = Q.maxm * 64;
    {
        // every load of the copy is issued before the first LDS store
        constexpr int NCP = (NR * 64 + 64 * JW - 1) / (64 * JW);
        double tm[NCP], tc[NCP], tp[NCP];
#pragma unroll
        for (int u = 0; u < NCP; ++u) {
            const int i = tid + u * 64 * JW;
            const int ic = i < lenf ? i : 0;
            tm[u] = Q.am[ic];
            tc[u] = Q.ac[ic];
            tp[u] = Q.ap[ic];
        }
#pragma unroll
        for (int u = 0; u < NCP; ++u) {
            const int i = tid + u * 64 * JW;
            if (i < NR * 64) {
                const bool ok = i < lenf;
                cam[i] = ok ? tm[u] : 0.0;
                cac[i] = ok ? tc[u] : 0.0;
                cap[i] = ok ? tp[u] : 0.0;
            }
        }
    }
    const int p = lane;
    const bool own = p < nch;
    const int m = own ? Q.cs[p + 1] - Q.cs[p] : 1;
    const int off = MAXR - (m - 1);
    int slot[NR];
    line_map<NR>(slot, Q.cs, n, nch, stride, lane);
    double pre[NR];
    const int nlines = F.n3m;
    const int step = (int)vb.gx * JW;
    const int iters = (nlines + step - 1) / step;
    int line = (int)vb.x * JW + wave;
    line_fetch<NR>(pre, F.hmn + (size_t)(line < nlines ? line : 0) * F.P, n, lane);
    __syncthreads();
    double* row = stage + p * stride;
    for (int it = 0; it < iters; ++it, line += step) {
        const bool act = line < nlines;
        line_put<NR>(stage, pre, slot);
        const int nxt = line + step;
        line_fetch<NR>(pre, F.hmn + (size_t)(nxt < nlines ? nxt : 0) * F.P, n, lane);
        wave_lds_sync();
        const int mode = line + 1;
        const double wrm = act ? F.wr[mode] : 0.0;
        const bool pin = act && (mode == F.null_mode);
        double cp[MAXR], vp[MAXR], y[MAXR];
        double cpp = 0.0, yl = 0.0, vl = 1.0, gg = 1.0, ys = 0.0, vs = 0.0, ws = 0.0, wl = 0.0;
        double alpha = 0.0, beta = 1.0, gamma = 0.0, delta = 0.0, ce = 0.0;
        double glast = 1.0;
#pragma unroll
        for (int tt = 0; tt < MAXR; ++tt) {
            const int t = tt - off;
            const bool real = t >= 0;
            const int ix = (real ? t : 0) * 64 + p;
            const double a = real ? cam[ix] : 0.0;
            const double b = real ? cac[ix] + wrm : 1.0;
            const double c = real ? cap[ix] : 0.0;
            const double d = real ? row[t] : 0.0;
            const double iv = 1.0 / (b - a * cpp);
            cpp = c * iv;
            yl = (d - a * yl) * iv;
            const double vn = (-a * vl) * iv;
            vl = real ? vn : vl;
            cp[tt] = cpp;
            vp[tt] = vl;
            y[tt] = yl;
            ys = ys + gg * yl;
            vs = real ? vs + gg * vl : vs;
            if (tt == MAXR - 1 && real) {
                wl = (-c) * iv;
                glast = gg;
            }
            gg = real ? -cpp * gg : gg;
        }
        if (own) {
            ws = glast * wl;
            const int ix = (m - 1) * 64 + p;
            const double ae = cam[ix];
            ce = cap[ix];
            alpha = ae * vl;
            beta = (cac[ix] + wrm) + ae * wl;
            delta = row[m - 1] - ae * yl;
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
        double x = 0.0;
#pragma unroll
        for (int tt = MAXR - 1; tt >= 0; --tt) {
            const int t = tt - off;
            if (tt == MAXR - 1) x = y[tt] + vp[tt] * Xm + wl * X;
            else x = (y[tt] + vp[tt] * Xm) - cp[tt] * x;
            if (t >= 0 && own) row[t] = x;
        }
        if (own) row[m - 1] = X;
        wave_lds_sync();
        if (act) line_get<NR>(F.hmn + (size_t)line * F.P, stage, slot, lane);
        wave_lds_sync();
    }
}
template <int MAXR, int JW>
__global__ __launch_bounds__(64 * JW) void k_jsolve_p(const FluidDev F)
{
    jsolve_p_body<MAXR, JW>(F, vb_hw());
}
template <int MAXR, int JW>
__global__ __launch_bounds__(64 * JW) void kb_jsolve_p(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    jsolve_p_body<MAXR, JW>(*Fp, vb);
}

template <int MAXR>
static void launch_h(const FluidDev& F, hipStream_t st)
{
    const int lines = 2 * F.n3m - 1;
    const int sd = F.jstride;
    if (jsolve_h_lds_bytes(MAXR, 8, sd, true) <= (size_t)F.lds_max) {
        const int g = std::max(2, std::min(256, (lines + 7) / 8) & ~1);
        hipLaunchKernelGGL((k_jsolve_hp<MAXR, 8>), dim3(g), dim3(512), jsolve_h_lds_bytes(MAXR, 8, sd, true), st, F);
    } else if (jsolve_h_lds_bytes(MAXR, 4, sd, true) <= (size_t)F.lds_max) {
        const int g = std::max(2, std::min(256, (lines + 3) / 4) & ~1);
        hipLaunchKernelGGL((k_jsolve_hp<MAXR, 4>), dim3(g), dim3(256), jsolve_h_lds_bytes(MAXR, 4, sd, true), st, F);
    } else if (jsolve_h_lds_bytes(MAXR, 2, sd, true) <= (size_t)F.lds_max) {
        const int g = std::max(2, std::min(512, (lines + 1) / 2) & ~1);
        hipLaunchKernelGGL((k_jsolve_hp<MAXR, 2>), dim3(g), dim3(128), jsolve_h_lds_bytes(MAXR, 2, sd, true), st, F);
    } else {
        const int g = std::max(2, std::min(2048, lines) & ~1);
        hipLaunchKernelGGL((k_jsolve_h<MAXR>), dim3(g), dim3(64), jsolve_h_lds_bytes(MAXR, 1, sd, false), st, F);
    }
}

template <int MAXR>
static void launch_p(const FluidDev& F, hipStream_t st)
{
    // the per-mode factors live in registers: 8 waves only for short chunks
    const int lines = F.n3m;
    if (MAXR <= 8 && jsolve_p_lds_bytes(MAXR, 8, F.jstride) <= (size_t)F.lds_max) {
        const int g = std::min(256, (lines + 7) / 8);
        hipLaunchKernelGGL((k_jsolve_p<MAXR, 8>), dim3(g), dim3(512), jsolve_p_lds_bytes(MAXR, 8, F.jstride), st, F);
    } else if (jsolve_p_lds_bytes(MAXR, 4, F.jstride) <= (size_t)F.lds_max) {
        const int g = std::min(256, (lines + 3) / 4);
        hipLaunchKernelGGL((k_jsolve_p<MAXR, 4>), dim3(g), dim3(256), jsolve_p_lds_bytes(MAXR, 4, F.jstride), st, F);
    } else {
        const int g = std::min(1024, lines);
        hipLaunchKernelGGL((k_jsolve_p<MAXR, 1>), dim3(g), dim3(64), jsolve_p_lds_bytes(MAXR, 1, F.jstride), st, F);
    }
}

static int jsolve_maxr(int maxm)
{
    const int mr = maxm - 1;
    return mr <= 8 ? 8 : mr <= 16 ? 16 : mr <= 24 ? 24 : mr <= 32 ? 32 : 64;
}

// Staging stride of the j-line kernels: every chunk (and, for the Helmholtz
// lines, the front padding up to MAXR plus the separator) fits; odd, so the
// lanes' rows fall in distinct banks.
int jsolve_stride(const FluidDev& F)
{
    const int mm = std::max(std::max(F.js[0].maxm, F.js[1].maxm), F.ps.maxm);
    return std::max(mm, jsolve_maxr(std::max(F.js[0].maxm, F.js[1].maxm)) + 1) | 1;
}

// Smallest LDS the j-line kernels need for this grid (their fallback shapes).
size_t jsolve_lds_bytes(const FluidDev& F)
{
    const int mh = jsolve_maxr(std::max(F.js[0].maxm, F.js[1].maxm));
    const int mp = jsolve_maxr(F.ps.maxm);
    return std::max(jsolve_h_lds_bytes(mh, 1, F.jstride, false), jsolve_p_lds_bytes(mp, 1, F.jstride));
}

void launch_jsolve_helm(const FluidDev& F, hipStream_t st)
{
    switch (jsolve_maxr(std::max(F.js[0].maxm, F.js[1].maxm))) {
    case 8: launch_h<8>(F, st); break;
    case 16: launch_h<16>(F, st); break;
    case 24: launch_h<24>(F, st); break;
    case 32: launch_h<32>(F, st); break;
    default: launch_h<64>(F, st); break;
    }
}

void launch_jsolve_poisson(const FluidDev& F, hipStream_t st)
{
    switch (jsolve_maxr(F.ps.maxm)) {
    case 8: launch_p<8>(F, st); break;
    case 16: launch_p<16>(F, st); break;
    case 24: launch_p<24>(F, st); break;
    case 32: launch_p<32>(F, st); break;
    default: launch_p<64>(F, st); break;
    }
}

// Batched j-lines: one kernel shape for the whole batch, the largest
// chunk-length class of its instances; the instances' LDS strides (set by
// the batch for that class) give the largest staging size.
template <int MAXR>
static void launch_h_b(BatchSet& B, hipStream_t st)
{
    int sd = 1;
    for (const auto& F : B.hF) sd = std::max(sd, F.jstride);
    const size_t lim = (size_t)B.hF[0].lds_max;
    auto lines = [](const FluidDev& F) { return 2 * F.n3m - 1; };
    if (jsolve_h_lds_bytes(MAXR, 8, sd, true) <= lim) {
        // FRIBO_JB_LPW lines per wave: each workgroup first stages the
        // factors of its instance in LDS, which many short lines amortise
        constexpr int LPW = FRIBO_JB_LPW;
        blaunch(B, BK_JH, bdims(B, [&](const FluidDev& F) { return dim3(std::max(2, std::min(256, (lines(F) + 8 * LPW - 1) / (8 * LPW)) & ~1)); }),
                kb_jsolve_hp<MAXR, 8>, dim3(512), jsolve_h_lds_bytes(MAXR, 8, sd, true), st);
    } else if (jsolve_h_lds_bytes(MAXR, 4, sd, true) <= lim) {
        blaunch(B, BK_JH, bdims(B, [&](const FluidDev& F) { return dim3(std::max(2, std::min(256, (lines(F) + 3) / 4) & ~1)); }),
                kb_jsolve_hp<MAXR, 4>, dim3(256), jsolve_h_lds_bytes(MAXR, 4, sd, true), st);
    } else if (jsolve_h_lds_bytes(MAXR, 2, sd, true) <= lim) {
        blaunch(B, BK_JH, bdims(B, [&](const FluidDev& F) { return dim3(std::max(2, std::min(512, (lines(F) + 1) / 2) & ~1)); }),
                kb_jsolve_hp<MAXR, 2>, dim3(128), jsolve_h_lds_bytes(MAXR, 2, sd, true), st);
    } else {
        blaunch(B, BK_JH, bdims(B, [&](const FluidDev& F) { return dim3(std::max(2, std::min(2048, lines(F)) & ~1)); }),
                kb_jsolve_h<MAXR>, dim3(64), jsolve_h_lds_bytes(MAXR, 1, sd, false), st);
    }
}

template <int MAXR>
static void launch_p_b(BatchSet& B, hipStream_t st)
{
    int sd = 1;
    for (const auto& F : B.hF) sd = std::max(sd, F.jstride);
    const size_t lim = (size_t)B.hF[0].lds_max;
    if (MAXR <= 8 && jsolve_p_lds_bytes(MAXR, 8, sd) <= lim) {
        blaunch(B, BK_JP, bdims(B, [](const FluidDev& F) { return dim3(std::min(256, (F.n3m + 7) / 8)); }),
                kb_jsolve_p<MAXR, 8>, dim3(512), jsolve_p_lds_bytes(MAXR, 8, sd), st);
    } else if (jsolve_p_lds_bytes(MAXR, 4, sd) <= lim) {
        blaunch(B, BK_JP, bdims(B, [](const FluidDev& F) { return dim3(std::min(256, (F.n3m + 3) / 4)); }),
                kb_jsolve_p<MAXR, 4>, dim3(256), jsolve_p_lds_bytes(MAXR, 4, sd), st);
    } else {
        blaunch(B, BK_JP, bdims(B, [](const FluidDev& F) { return dim3(std::min(1024, F.n3m)); }),
                kb_jsolve_p<MAXR, 1>, dim3(64), jsolve_p_lds_bytes(MAXR, 1, sd), st);
    }
}

// Chunk-length classes of the j-line kernels: Helmholtz lines, Poisson lines.
int jsolve_class_h(const FluidDev& F) { return jsolve_maxr(std::max(F.js[0].maxm, F.js[1].maxm)); }
int jsolve_class_p(const FluidDev& F) { return jsolve_maxr(F.ps.maxm); }

// LDS stride of an instance's j-lines when the batch runs chunk class mh
// (Helmholtz) for every instance.
int jsolve_stride_for(const FluidDev& F, int mh)
{
    const int mm = std::max(std::max(F.js[0].maxm, F.js[1].maxm), F.ps.maxm);
    return std::max(mm, mh + 1) | 1;
}

void launch_jsolve_helm_b(BatchSet& B, hipStream_t st)
{
    int c = 8;
    for (const auto& F : B.hF) c = std::max(c, jsolve_class_h(F));
    switch (c) {
    case 8: launch_h_b<8>(B, st); break;
    case 16: launch_h_b<16>(B, st); break;
    case 24: launch_h_b<24>(B, st); break;
    case 32: launch_h_b<32>(B, st); break;
    default: launch_h_b<64>(B, st); break;
    }
}

void launch_jsolve_poisson_b(BatchSet& B, hipStream_t st)
{
    int c = 8;
    for (const auto& F : B.hF) c = std::max(c, jsolve_class_p(F));
    switch (c) {
    case 8: launch_p_b<8>(B, st); break;
    case 16: launch_p_b<16>(B, st); break;
    case 24: launch_p_b<24>(B, st); break;
    case 32: launch_p_b<32>(B, st); break;
    default: launch_p_b<64>(B, st); break;
    }
}

// ---------------------------------------------------------------------------
// k-direction solves.  Field f = 0: q2 (rows kc=1..n3m), f = 1: q3 (interior
// rows kc=2..n3m of the n3 system; the Dirichlet rows dqb3s/dqb3n enter the
// first/last rhs).  The chunk summaries of column j live at
// ((f*n2m + j-1)*64 + p): chunk-fastest, so the separator pass reads a
// column's 64 chunks with one coalesced access.

__device__ inline size_t kidx(const FluidDev& F, int f, int j, int p)
{
    return ((size_t)f * F.n2m + (size_t)(j - 1)) * FRIBO_KMAXCH + p;
}

// Boundary increments of the first (b0) and last (bn) row of a k-line
// (solve_fluid.F90:130-131,152-153 for q2; :215,223 for q3).
__device__ inline void kbound(const FluidDev& F, const KSys& K, int f, int j, const DevScalars* S, double& b0,
                              double& bn)
{
    if (f == 0) {
        b0 = F.dqb2s[j] * S->ugkks;
        bn = F.dqb2n[j] * S->ugkkn;
    } else {
        b0 = -(K.bc[0] * F.dqb3s[j]);
        bn = -(K.bc[1] * F.dqb3n[j]);
    }
}

// Row coefficients of the interior rows s..s+m-1 of a chunk, one row per
// lane (lane l holds row s + l + 64g in register g; clamped past the end):
// every lane of a workgroup works on the same chunk, so each step takes its
// row from a lane instead of issuing a scalar load it must wait for.
template <int NG>
__device__ inline void chunk_rows(const double* __restrict__ a, int s, int m, double (&v)[NG])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int r = lane + 64 * g;
        v[g] = a[s + (r < m ? r : (m > 0 ? m - 1 : 0))];
    }
}

template <int NG>
__device__ inline double chunk_row(const double (&v)[NG], int t)
{
    return lane_bcast(NG == 1 || t < 64 ? v[0] : v[NG - 1], t & 63);
}

// Pass 1: thread (j, chunk p) eliminates the interior rows of its chunk and
// leaves the chunk's separator data (sum g*y', last y', separator rhs).  All
// loads of the chunk are issued together (clamped rows past the chunk end);
// lanes past the last column load column n2m and store nothing.
template <int MAXM>
__device__ __forceinline__ void ksolve1_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int NG = (MAXM + 63) / 64;
    static_assert(NG <= 2, "chunk rows");
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = vb.z, p = vb.y;
    const KSys& K = F.ks[f];
    if (p >= K.nch) return;
    const int j = vb.x * 64 + threadIdx.x + 1;
    const bool act = j <= F.n2m;
    const int jc = act ? j : F.n2m;
    const double* __restrict__ src = (f == 0 ? F.sA : F.sB) + fidx(F, jc, K.kofs);
    const size_t P = F.P;
    const int s = K.cs[p], e = K.cs[p + 1] - 1, m = e - s;
    double cS[NG], ca[NG], ci[NG], cg[NG];
    chunk_rows<NG>(K.S, s, m, cS);
    chunk_rows<NG>(K.a, s, m, ca);
    chunk_rows<NG>(K.inv, s, m, ci);
    chunk_rows<NG>(K.g, s, m, cg);
    double b0, bn;
    kbound(F, K, f, jc, S, b0, bn);
    // all loads of the chunk first (one memory round trip), then the recurrence
    double dv[MAXM];
#pragma unroll
    for (int t = 0; t < MAXM; ++t) dv[t] = src[(size_t)(s + (t < m ? t : 0)) * P];
    const double dsep = src[(size_t)e * P];
    double yp = 0.0, acc = 0.0;
#pragma unroll
    for (int t = 0; t < MAXM; ++t) {
        if (t < m) {
            const int i = s + t;
            double d = dv[t] * chunk_row<NG>(cS, t);
            d = (i == 0) ? d + b0 : d;
            yp = (d - chunk_row<NG>(ca, t) * yp) * chunk_row<NG>(ci, t);
            acc = acc + chunk_row<NG>(cg, t) * yp;
        }
    }
    double de = dsep * K.S[e];
    de = (e == 0) ? de + b0 : de;
    de = (e == K.n - 1) ? de + bn : de;
    if (!act) return;
    const size_t o = kidx(F, f, j, p);
    F.kYs[o] = acc;
    F.kYe[o] = yp;
    F.kDe[o] = de;
}
template <int MAXM>
__global__ __launch_bounds__(64) void k_ksolve1(const FluidDev F)
{
    ksolve1_body<MAXM>(F, vb_hw());
}
template <int MAXM>
__global__ __launch_bounds__(64) void kb_ksolve1(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ksolve1_body<MAXM>(*Fp, vb);
}

// Pass 2: one wave per column solves the nch separators (PCR over lanes with
// the precomputed line-invariant multipliers).
__device__ __forceinline__ void ksolve2_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = vb.y, j = vb.x + 1, p = threadIdx.x;
    const KSys& K = F.ks[f];
    const size_t o = kidx(F, f, j, p);
    double delta = 0.0;
    if (p < K.nch) {
        const int e = K.cs[p + 1] - 1;
        const double ysn = (p + 1 < K.nch) ? F.kYs[o + 1] : 0.0;
        delta = F.kDe[o] - K.a[e] * F.kYe[o] - K.c[e] * ysn;
    }
    int li = 0;
    for (int h = 1; h < 64; h <<= 1, ++li) {
        double dm = __shfl_up(delta, h, 64), dp = __shfl_down(delta, h, 64);
        if (p - h < 0) dm = 0.0;
        if (p + h >= 64) dp = 0.0;
        delta = delta - K.k1[li * 64 + p] * dm - K.k2[li * 64 + p] * dp;
    }
    if (p < K.nch) F.kX[o] = delta * K.ib[p];
}
__global__ __launch_bounds__(64) void k_ksolve2(const FluidDev F)
{
    ksolve2_body(F, vb_hw());
}
__global__ __launch_bounds__(64) void kb_ksolve2(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ksolve2_body(*Fp, vb);
}

// Pass 3: chunk back-substitution fused with q += dq (solve_fluid.F90:246-250).
template <int MAXM>
__device__ __forceinline__ void ksolve3_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int NG = (MAXM + 63) / 64;
    static_assert(NG <= 2, "chunk rows");
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = vb.z, p = vb.y;
    const KSys& K = F.ks[f];
    if (p >= K.nch) return;
    const int j = vb.x * 64 + threadIdx.x + 1;
    const bool act = j <= F.n2m;
    const int jc = act ? j : F.n2m;
    const size_t P = F.P;
    const double* __restrict__ src = (f == 0 ? F.sA : F.sB) + fidx(F, jc, K.kofs);
    double* __restrict__ q = (f == 0 ? F.q2 : F.q3) + fidx(F, jc, K.kofs);
    const int s = K.cs[p], e = K.cs[p + 1] - 1, m = e - s;   // m interior rows
    double cS[NG], ca[NG], ci[NG], cv[NG], cc[NG];
    chunk_rows<NG>(K.S, s, m, cS);
    chunk_rows<NG>(K.a, s, m, ca);
    chunk_rows<NG>(K.inv, s, m, ci);
    chunk_rows<NG>(K.vp, s, m, cv);
    chunk_rows<NG>(K.cp, s, m, cc);
    const size_t o = kidx(F, f, jc, p);
    const double X = F.kX[o];
    const double Xm = (p > 0) ? F.kX[o - 1] : 0.0;
    const double wlp = K.wl[p];
    const bool zero_row = (f == 0 && j == 1);
    double b0, bn;
    kbound(F, K, f, jc, S, b0, bn);
    // all loads of the chunk first (one memory round trip), then the recurrence
    double y[MAXM], qv[MAXM];
#pragma unroll
    for (int t = 0; t < MAXM; ++t) {
        const int i = s + (t < m ? t : 0);
        y[t] = src[(size_t)i * P];
        qv[t] = q[(size_t)i * P];
    }
    const double qe = q[(size_t)e * P];
    double yp = 0.0;
#pragma unroll
    for (int t = 0; t < MAXM; ++t) {
        if (t < m) {
            const int i = s + t;
            double d = y[t] * chunk_row<NG>(cS, t);
            d = (i == 0) ? d + b0 : d;
            yp = (d - chunk_row<NG>(ca, t) * yp) * chunk_row<NG>(ci, t);
            y[t] = yp;
        }
    }
    // every lane stays active through the broadcasts; only stores are masked
    double x = 0.0;
    double* qn2 = (f == 0 ? F.q2 : F.q3) + fidx(F, F.n2, K.kofs);
#pragma unroll
    for (int t = MAXM - 1; t >= 0; --t) {
        if (t < m) {
            const int i = s + t;
            if (t == m - 1) x = y[t] + chunk_row<NG>(cv, t) * Xm + wlp * X;
            else x = (y[t] + chunk_row<NG>(cv, t) * Xm) - chunk_row<NG>(cc, t) * x;
            if (act) q[(size_t)i * P] = zero_row ? 0.0 : qv[t] + x;
            if (act && zero_row) qn2[(size_t)i * P] = 0.0;
        }
    }
    if (!act) return;
    q[(size_t)e * P] = zero_row ? 0.0 : qe + X;
    if (zero_row) qn2[(size_t)e * P] = 0.0;
    if (f == 1) {
        double* qf = F.q3;
        if (p == 0) {
            const size_t c = fidx(F, j, 1);
            qf[c] = qf[c] + F.dqb3s[j];
        }
        if (p == K.nch - 1) {
            const size_t c = fidx(F, j, F.n3);
            qf[c] = qf[c] + F.dqb3n[j];
        }
    }
}
template <int MAXM>
__global__ __launch_bounds__(64) void k_ksolve3(const FluidDev F)
{
    ksolve3_body<MAXM>(F, vb_hw());
}
template <int MAXM>
__global__ __launch_bounds__(64) void kb_ksolve3(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ksolve3_body<MAXM>(*Fp, vb);
}

// One-pass form of the three passes above for chunks of at most 32 rows
// (lines of up to 2048 rows): a workgroup of 1024 threads owns KF_COLS = 16
// columns of one field and every chunk of them; thread (column jl, chunk p)
// with jl = lane % 16, p = 4 * wave + lane / 16, so a wave's loads cover 16
// columns (128 bytes) of 4 chunks.  The chunk's forward-eliminated rows stay
// in registers across the separator solve, so the right-hand side is read
// once: read src, read q, write q.  The 16 lanes of a chunk hold its
// coefficient rows (lane u rows u and u + 16) and each elimination step takes
// its row by a DPP broadcast within the 16-lane row (no LDS traffic); the
// chunk summaries go to LDS, wave w solves the separator system of column w
// with the multipliers and operation order of k_ksolve2, and the
// back-substitution is that of k_ksolve3, so both forms give identical
// results.
constexpr int KF_COLS = 16;

// Lane n of each 16-lane row of the wave, broadcast to the row (DPP
// row_newbcast), for a double.
template <int N>
__device__ inline double row_bcast(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + N, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + N, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Row T of the chunk held by the lanes of this 16-lane row (T < 16 * NG).
template <int T, int NG>
__device__ inline double krow(const double (&v)[NG])
{
    return row_bcast<T & 15>(v[T >> 4]);
}

template <int... I, class Fn>
__device__ inline void kf_unroll(std::integer_sequence<int, I...>, Fn fn)
{
    (fn(std::integral_constant<int, I>{}), ...);
}

// Raw buffer access to a field: a uniform descriptor plus one 32-bit byte
// offset per lane (a 64-bit address would take two registers per row).
typedef unsigned int kf_u32x2 __attribute__((ext_vector_type(2)));
__device__ inline __amdgpu_buffer_rsrc_t kf_rsrc(const double* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), 0, 0x7fffffff, 0x00020000);
}
// voff: per-lane byte offset; soff: uniform byte offset (a scalar register)
__device__ inline double kf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0)
{
    const kf_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return __hiloint2double((int)v.y, (int)v.x);
}
__device__ inline void kf_store(__amdgpu_buffer_rsrc_t r, int voff, double d, int soff = 0)
{
    kf_u32x2 v;
    v.x = (unsigned)__double2loint(d);
    v.y = (unsigned)__double2hiint(d);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
}

template <int MAXM>
__device__ __forceinline__ void ksolve_fused_body(const FluidDev& F, const VBlock& vb)
{
    static_assert(MAXM <= 32, "fused k-solve: chunks of at most 32 rows");
    constexpr int MI = MAXM - 1;                  // interior rows of the longest chunk
    constexpr int NG = (MI + 15) / 16;            // coefficient rows per lane
    __shared__ double sY[3][64][KF_COLS + 1];
    __shared__ double sX[64][KF_COLS + 1];
    const DevScalars* S = F.S;
    if (S->err) return;
    const int f = vb.z;
    const KSys& K = F.ks[f];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int jl = lane & (KF_COLS - 1), p = 4 * wave + (lane >> 4);
    const int nch = K.nch;
    const int j = vb.x * KF_COLS + jl + 1;
    const bool act = j <= F.n2m;
    const int jc = act ? j : F.n2m;
    // byte offsets of a field fit in 31 bits (the launcher checks), so every
    // access is a uniform buffer descriptor plus a 32-bit lane offset
    const int P8 = F.P * 8;
    const int col = ((jc - 1) + (K.kofs + 2) * F.P) * 8;  // row 0 of this column's line
    const bool own = p < nch;
    const int pc = own ? p : nch - 1;             // idle rows of lanes work on the last chunk, store nothing
    const int s = K.cs[pc], e = K.cs[pc + 1] - 1, m = e - s;
    const __amdgpu_buffer_rsrc_t src = kf_rsrc(f == 0 ? F.sA : F.sB);
    // row tt of the chunk: lane offset of row s plus the uniform tt * P8
    const int v0 = col + s * P8;
    double y[MI];
#pragma unroll
    for (int tt = 0; tt < MI; ++tt) y[tt] = tt < m ? kf_load(src, v0, tt * P8) : 0.0;
    const double dsep = kf_load(src, col + e * P8);
    // coefficient rows of the chunk: lane u of the row holds rows u + 16 g
    double cS[NG], ca[NG], ci[NG], cg[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int r = s + min(jl + 16 * g, m > 0 ? m - 1 : 0);
        cS[g] = K.S[r];
        ca[g] = K.a[r];
        ci[g] = K.inv[r];
        cg[g] = K.g[r];
    }
    const double Se = K.S[e];
    double b0, bn;
    kbound(F, K, f, jc, S, b0, bn);
    double yp = 0.0, acc = 0.0;
    kf_unroll(std::make_integer_sequence<int, MI>{}, [&](auto ic) {
        constexpr int tt = decltype(ic)::value;
        // one step at a time: hoisted, the broadcasts of all steps would be
        // live at once and spill
        __builtin_amdgcn_sched_barrier(0);
        const double rS = krow<tt>(cS), ra = krow<tt>(ca), ri = krow<tt>(ci), rg = krow<tt>(cg);
        if (tt < m) {
            const int i = s + tt;
            double d = y[tt] * rS;
            d = (i == 0) ? d + b0 : d;
            yp = (d - ra * yp) * ri;
            acc = acc + rg * yp;
            y[tt] = yp;
        }
    });
    double de = dsep * Se;
    de = (e == 0) ? de + b0 : de;
    de = (e == K.n - 1) ? de + bn : de;
    if (own) {
        sY[0][p][jl] = acc;
        sY[1][p][jl] = yp;
        sY[2][p][jl] = de;
    }
    // back-substitution rows and the separator row of q, loaded while the
    // separator systems are solved
    const __amdgpu_buffer_rsrc_t q = kf_rsrc(f == 0 ? F.q2 : F.q3);
    const double qe = kf_load(q, col + e * P8);
    double cv[NG], cc[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int r = s + min(jl + 16 * g, m > 0 ? m - 1 : 0);
        cv[g] = K.vp[r];
        cc[g] = K.cp[r];
    }
    __syncthreads();
    {
        // separator system of column wave (k_ksolve2's operations); lane = chunk
        const int c = wave, q = lane;
        double delta = 0.0;
        if (q < nch) {
            const int eq = K.cs[q + 1] - 1;
            const double ysn = (q + 1 < nch) ? sY[0][q + 1][c] : 0.0;
            delta = sY[2][q][c] - K.a[eq] * sY[1][q][c] - K.c[eq] * ysn;
        }
        int li = 0;
        for (int h = 1; h < 64; h <<= 1, ++li) {
            double dm = __shfl_up(delta, h, 64), dp = __shfl_down(delta, h, 64);
            if (q - h < 0) dm = 0.0;
            if (q + h >= 64) dp = 0.0;
            delta = delta - K.k1[li * 64 + q] * dm - K.k2[li * 64 + q] * dp;
        }
        if (q < nch) sX[q][c] = delta * K.ib[q];
    }
    __syncthreads();
    const double X = sX[pc][jl];
    const double Xm = (pc > 0) ? sX[pc - 1][jl] : 0.0;
    const double wlp = K.wl[pc];
    const bool st = own && act;
    const bool zero_row = (f == 0 && j == 1);
    // back-substitution into y (x of every interior row, no memory traffic)
    double x = 0.0;
    kf_unroll(std::make_integer_sequence<int, MI>{}, [&](auto ic) {
        constexpr int tt = MI - 1 - decltype(ic)::value;
        __builtin_amdgcn_sched_barrier(0);
        const double rv = krow<tt>(cv), rc = krow<tt>(cc);
        if (tt < m) {
            if (tt == m - 1) x = y[tt] + rv * Xm + wlp * X;
            else x = (y[tt] + rv * Xm) - rc * x;
            y[tt] = x;
        }
    });
    // q += x, rows in batches whose loads are issued together (the stores of
    // a batch precede the next batch's loads, which the compiler keeps in
    // order since both address the same field)
    const int coln = ((F.n2 - 1) + (K.kofs + 2) * F.P) * 8;   // column n2
    constexpr int QB = MI < 8 ? MI : 8;
#pragma unroll
    for (int b = 0; b < MI; b += QB) {
        double qv[QB];
#pragma unroll
        for (int u = 0; u < QB; ++u) {
            const int tt = b + u;
            if (tt < MI) qv[u] = tt < m ? kf_load(q, v0, tt * P8) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < QB; ++u) {
            const int tt = b + u;
            if (tt < MI && tt < m && st) kf_store(q, v0, zero_row ? 0.0 : qv[u] + y[tt], tt * P8);
        }
    }
    if (st && zero_row)
        for (int i = s; i < e; ++i) kf_store(q, coln + i * P8, 0.0);
    if (!st) return;
    kf_store(q, col + e * P8, zero_row ? 0.0 : qe + X);
    if (zero_row) kf_store(q, coln + e * P8, 0.0);
    if (f == 1) {
        double* qf = F.q3;
        if (p == 0) {
            const size_t c0 = fidx(F, j, 1);
            qf[c0] = qf[c0] + F.dqb3s[j];
        }
        if (p == nch - 1) {
            const size_t c1 = fidx(F, j, F.n3);
            qf[c1] = qf[c1] + F.dqb3n[j];
        }
    }
}
template <int MAXM>
__global__ __launch_bounds__(1024) void k_ksolve_fused(const FluidDev F)
{
    ksolve_fused_body<MAXM>(F, vb_hw());
}
template <int MAXM>
__global__ __launch_bounds__(1024) void kb_ksolve_fused(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    ksolve_fused_body<MAXM>(*Fp, vb);
}

// Largest chunk (rows incl. the separator) of either k-system.
static inline int kmaxm(const FluidDev& F) { return std::max(F.ks[0].maxm, F.ks[1].maxm); }

bool ksolve_fused_fits(const FluidDev& F)
{
    return kmaxm(F) <= 32 && (long)F.P * F.nk * 8 < (1L << 31);
}

void launch_ksolve(const FluidDev& F, hipStream_t st)
{
    int nchmax = F.ks[0].nch > F.ks[1].nch ? F.ks[0].nch : F.ks[1].nch;
    int maxm = kmaxm(F);
    if (F.kform && ksolve_fused_fits(F)) {
        const dim3 gf((F.n2m + KF_COLS - 1) / KF_COLS, 1, 2);
        if (maxm <= 8) hipLaunchKernelGGL(k_ksolve_fused<8>, gf, dim3(1024), 0, st, F);
        else if (maxm <= 16) hipLaunchKernelGGL(k_ksolve_fused<16>, gf, dim3(1024), 0, st, F);
        else hipLaunchKernelGGL(k_ksolve_fused<32>, gf, dim3(1024), 0, st, F);
        return;
    }
    dim3 g((F.n2m + 63) / 64, nchmax, 2);
    // maxm counts the separator row; the passes loop over the m-1 interior rows
    if (maxm - 1 <= 8) hipLaunchKernelGGL(k_ksolve1<8>, g, dim3(64), 0, st, F);
    else if (maxm - 1 <= 16) hipLaunchKernelGGL(k_ksolve1<16>, g, dim3(64), 0, st, F);
    else if (maxm - 1 <= 32) hipLaunchKernelGGL(k_ksolve1<32>, g, dim3(64), 0, st, F);
    else if (maxm - 1 <= 64) hipLaunchKernelGGL(k_ksolve1<64>, g, dim3(64), 0, st, F);
    else hipLaunchKernelGGL(k_ksolve1<128>, g, dim3(64), 0, st, F);
    hipLaunchKernelGGL(k_ksolve2, dim3(F.n2m, 2), dim3(64), 0, st, F);
    if (maxm <= 8) hipLaunchKernelGGL(k_ksolve3<8>, g, dim3(64), 0, st, F);
    else if (maxm <= 16) hipLaunchKernelGGL(k_ksolve3<16>, g, dim3(64), 0, st, F);
    else if (maxm <= 32) hipLaunchKernelGGL(k_ksolve3<32>, g, dim3(64), 0, st, F);
    else if (maxm <= 64) hipLaunchKernelGGL(k_ksolve3<64>, g, dim3(64), 0, st, F);
    else hipLaunchKernelGGL(k_ksolve3<128>, g, dim3(64), 0, st, F);
}

void launch_ksolve_b(BatchSet& B, hipStream_t st)
{
    int maxm = 0;
    bool fused = true;
    for (const auto& F : B.hF) {
        maxm = std::max(maxm, kmaxm(F));
        fused = fused && F.kform && ksolve_fused_fits(F);
    }
    if (fused) {
        const auto gf = bdims(B, [](const FluidDev& F) { return dim3((F.n2m + KF_COLS - 1) / KF_COLS, 1, 2); });
        if (maxm <= 8) blaunch(B, BK_KF, gf, kb_ksolve_fused<8>, dim3(1024), 0, st);
        else if (maxm <= 16) blaunch(B, BK_KF, gf, kb_ksolve_fused<16>, dim3(1024), 0, st);
        else blaunch(B, BK_KF, gf, kb_ksolve_fused<32>, dim3(1024), 0, st);
        return;
    }
    const auto g = bdims(B, [](const FluidDev& F) {
        return dim3((F.n2m + 63) / 64, std::max(F.ks[0].nch, F.ks[1].nch), 2);
    });
    if (maxm - 1 <= 8) blaunch(B, BK_K1, g, kb_ksolve1<8>, dim3(64), 0, st);
    else if (maxm - 1 <= 16) blaunch(B, BK_K1, g, kb_ksolve1<16>, dim3(64), 0, st);
    else if (maxm - 1 <= 32) blaunch(B, BK_K1, g, kb_ksolve1<32>, dim3(64), 0, st);
    else if (maxm - 1 <= 64) blaunch(B, BK_K1, g, kb_ksolve1<64>, dim3(64), 0, st);
    else blaunch(B, BK_K1, g, kb_ksolve1<128>, dim3(64), 0, st);
    blaunch(B, BK_K2, bdims(B, [](const FluidDev& F) { return dim3(F.n2m, 2); }), kb_ksolve2, dim3(64), 0, st);
    if (maxm <= 8) blaunch(B, BK_K3, g, kb_ksolve3<8>, dim3(64), 0, st);
    else if (maxm <= 16) blaunch(B, BK_K3, g, kb_ksolve3<16>, dim3(64), 0, st);
    else if (maxm <= 32) blaunch(B, BK_K3, g, kb_ksolve3<32>, dim3(64), 0, st);
    else if (maxm <= 64) blaunch(B, BK_K3, g, kb_ksolve3<64>, dim3(64), 0, st);
    else blaunch(B, BK_K3, g, kb_ksolve3<128>, dim3(64), 0, st);
}
