// Chunk elimination of the per-y-mode z-lines of the y-order Poisson solve
// (the shifted operator A_z + wy(j) of phcalc_outflow.F90:34-56 restated
// along z), shared by the line kernels of fribo_ysolve.hip and by the
// epilogue of the forward transform GEMM, which runs the first pass on its
// own output tile.
#pragma once
#include "fribo_internal.h"

// summary q of chunk p of y-mode j (1-based): chunk fastest
__host__ __device__ inline size_t ysum_idx(int n2m, int q, int j, int p)
{
    return ((size_t)q * n2m + (size_t)(j - 1)) * FRIBO_KMAXCH + p;
}

// Rows k = s+1 .. s+m of the z operator for the interior rows of chunk p,
// one row per lane (every lane of a wave works on the same chunk), taken by
// each step from its lane.
template <int NG>
__device__ inline void yrows(const double* __restrict__ a, int s, int m, double (&v)[NG])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int r = lane + 64 * g;
        v[g] = a[s + 1 + (r < m ? r : (m > 0 ? m - 1 : 0))];
    }
}

template <int NG>
__device__ inline double yrow(const double (&v)[NG], int t)
{
    return lane_bcast(NG == 1 || t < 64 ? v[0] : v[NG - 1], t & 63);
}

// Chunk elimination of one line, rows s..s+m-1 interior with right-hand
// sides dv[0..m-1] (the separator row s+m follows).  Returns the recurrence
// state; cp/vp/y are kept when STORE.  Every lane of the wave must take part
// (the row coefficients are broadcast from lanes).
template <int MAXM, bool STORE>
__device__ inline void ychunk_rows(const double* amphk, const double* acphk, const double* apphk, int s, int m,
                                   double w, const double (&dv)[MAXM], double* cp, double* vp, double* y,
                                   double& yl, double& vl, double& ys, double& vs, double& ws, double& wl)
{
    constexpr int NG = (MAXM + 63) / 64;
    static_assert(NG <= 2, "chunk rows");
    double ra[NG], rb[NG], rc[NG];
    yrows<NG>(amphk, s, m, ra);
    yrows<NG>(acphk, s, m, rb);
    yrows<NG>(apphk, s, m, rc);
    double cpp = 0.0, gg = 1.0, glast = 1.0;
    yl = 0.0;
    vl = 1.0;
    ys = 0.0;
    vs = 0.0;
    wl = 0.0;
#pragma unroll
    for (int t = 0; t < MAXM; ++t) {
        if (t < m) {
            const double a = yrow<NG>(ra, t), b = yrow<NG>(rb, t) + w, c = yrow<NG>(rc, t);
            const double iv = recip(b - a * cpp);
            cpp = c * iv;
            yl = (dv[t] - a * yl) * iv;
            vl = (-a * vl) * iv;
            if (STORE) {
                cp[t] = cpp;
                vp[t] = vl;
                y[t] = yl;
            }
            ys = ys + gg * yl;
            vs = vs + gg * vl;
            if (t == m - 1) {
                wl = (-c) * iv;
                glast = gg;
            }
            gg = -cpp * gg;
        }
    }
    ws = glast * wl;
}

// The six separator summaries of a chunk: e is the separator row, de its
// right-hand side.
__device__ inline void ysum_vals(double (&v)[6], const double* amphk, const double* acphk, int e, double w, double de,
                                 double yl, double vl, double wl, double ys, double vs, double ws)
{
    const int k = e + 1;
    const double ae = amphk[k];
    v[0] = ae * vl;                        // alpha
    v[1] = (acphk[k] + w) + ae * wl;       // beta before the coupling to chunk p+1
    v[2] = de - ae * yl;                   // delta before the coupling
    v[3] = ys;
    v[4] = vs;
    v[5] = ws;
}

// The six separator summaries of chunk p of y-mode j, stored.
__device__ inline void ysum_store(double* Y, int n2m, int j, int p, const double* amphk, const double* acphk, int e,
                                  double w, double de, double yl, double vl, double wl, double ys, double vs,
                                  double ws)
{
    double v[6];
    ysum_vals(v, amphk, acphk, e, w, de, yl, vl, wl, ys, vs, ws);
#pragma unroll
    for (int q = 0; q < 6; ++q) Y[ysum_idx(n2m, q, j, p)] = v[q];
}
