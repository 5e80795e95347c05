// Staggered-grid stencil kernels of the fluid step (gfx950).
//
//   k_cfl   cfl.F90:4-51                  -> per-block partial max / sum
//   k_bc    fluid_tsch.F90:34-77 (dt control, RK coefficients),
//           iobcs.F90:5-77 (boucdq), and the line-invariant factors of the
//           partitioned k-direction solves of solve_fluid.F90:107-255
//   k_rhs   HDNL.F90:5-171 + INVTR.F90:4-280 fused: one pass over q2,q3,pr,
//           ru2,ru3 writes ru2,ru3,rhs2,rhs3 (the convective terms dph/qcap
//           stay in registers)
//   k_div   divergence.F90:15-42
//   k_upd   updvp.F90:5-105 (updvp + prcalc) + partial sums of pr_submean
//   k_mean  updvp.F90:112-144 (pr_submean) + iobcs.F90:81-108 (boucqt)
//
// The arithmetic of every formula keeps the reference's operation order;
// the library is compiled with -ffp-contract=off so these kernels round
// exactly like the FP64 reference loops.  Deviations from the reference's
// sequential sums: the CFL/pressure-mean/outflow-flux reductions use a fixed
// tree order (deterministic run to run).
#include "fribo_batch.h"

#include <algorithm>


// Fixed-order reductions over the NT threads of a block: a butterfly inside
// each wave (every lane ends with the same value: the pairwise sums are
// symmetric), then the wave results in wave order.  Every thread returns the
// result.  sh holds NT/64 values and may be reused by consecutive calls.
template <int NT>
__device__ inline double block_sum(double* sh, double v)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    lds_barrier();
    if ((t & 63) == 0) sh[t >> 6] = v;
    lds_barrier();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r = r + sh[w];
    return r;
}

template <int NT>
__device__ inline double block_max(double* sh, double v)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    lds_barrier();
    if ((t & 63) == 0) sh[t >> 6] = v;
    lds_barrier();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r = fmax(r, sh[w]);
    return r;
}

// ---------------------------------------------------------------------------
// cfl.F90:19-38 -- partial max of the local CFL and partial volume sum of q3.
// Block b owns a contiguous range of planes k; thread t owns the columns
// j = t+1, t+257, ... and carries q3 of plane k+1 into the next plane, so q2
// and q3 are each read from HBM once.  Blocks without planes write the
// neutral partials (max seed 1e-8, sum 0).
// cfl.F90:19-38: workgroup (bx, by) takes columns 256*bx+1.. and planes
// FRIBO_CFL_KT*by+1..; every load of its planes is issued before the
// arithmetic (q3 of the plane above comes from the next plane's load, one
// extra plane at the top), and it leaves one partial (max, q3 sum).
__device__ __forceinline__ void cfl_body(const FluidDev& F, const VBlock& vb)
{
    constexpr int KT = FRIBO_CFL_KT;
    __shared__ double sh[256];
    const int t = threadIdx.x;
    const int j = (int)vb.x * 256 + t + 1;
    const int k0 = (int)vb.y * KT + 1;
    const int k1 = min(k0 + KT - 1, F.n3m);
    const bool act = j <= F.n2m;
    const int jc = act ? j : F.n2m;
    const double* __restrict__ q2 = F.q2;
    const double* __restrict__ q3 = F.q3;
    const size_t P = F.P, c0 = fidx(F, jc, k0);
    double a2[KT], b2[KT], a3[KT + 1];
#pragma unroll
    for (int u = 0; u < KT; ++u) {
        const size_t c = c0 + (size_t)(min(k0 + u, k1) - k0) * P;
        a2[u] = q2[c];
        b2[u] = q2[c + 1];
        a3[u] = q3[c];
    }
    a3[KT] = q3[c0 + (size_t)(k1 + 1 - k0) * P];
    const double u2 = F.udx2m[jc], dy = F.rc[jc + 1] - F.rc[jc];
    double m = 1.0e-8, s = 0.0;
#pragma unroll
    for (int u = 0; u < KT; ++u) {
        const int k = k0 + u;
        if (act && k <= k1) {
            const double u3 = F.udx3m[k], dz = F.zz[k + 1] - F.zz[k];
            const double q3b = (k < k1) ? a3[u + 1] : a3[KT];
            const double qcf = fabs((a2[u] + b2[u]) * 0.5 * u2) + fabs((a3[u] + q3b) * 0.5 * u3);
            m = fmax(m, qcf);
            s = s + a3[u] * dz * dy;
        }
    }
    const double bm = block_max<256>(sh, m);
    const double bs = block_sum<256>(sh, s);
    if (t == 0) {
        const int part = (int)vb.y * (int)vb.gx + (int)vb.x;
        F.cfl_part[2 * part] = bm;
        F.cfl_part[2 * part + 1] = bs;
    }
}
__global__ __launch_bounds__(256) void k_cfl(const FluidDev F)
{
    cfl_body(F, vb_hw());
}
__global__ __launch_bounds__(256) void kb_cfl(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    cfl_body(*Fp, vb);
}

static dim3 grid_cfl(const FluidDev& F)
{
    return dim3((F.n2m + 255) / 256, (F.n3m + FRIBO_CFL_KT - 1) / FRIBO_CFL_KT);
}

void launch_cfl(const FluidDev& F, hipStream_t st)
{
    hipLaunchKernelGGL(k_cfl, grid_cfl(F), dim3(256), 0, st, F);
}

void launch_cfl_b(BatchSet& B, hipStream_t st)
{
    blaunch(B, BK_CFL, bdims(B, grid_cfl), kb_cfl, dim3(256), 0, st);
}

// ---------------------------------------------------------------------------
// Step scalars shared by every workgroup of k_bc (fluid_tsch.F90:34-77): each
// workgroup reduces the CFL partials itself (a max, exact in any order) so the
// system-preparation workgroups need no inter-workgroup synchronisation.
struct StepScalars {
    double dt, cflm, q3sum, beta, al, ga, ro, betadx;
    int err;
};

__device__ StepScalars step_scalars(const FluidDev& F, int n, int ns, double dt_in, double dt_o, double s_dt,
                                    double s_cflm, double s_beta, int s_err, double* sh, double* sh2)
{
    __shared__ StepScalars out;
    const int t = threadIdx.x;
    double cmax = 0.0, q3s = 0.0;
    if (n == 1) {
        double m = 1.0e-8, s = 0.0;
        for (int i = t; i < F.ncfl; i += 1024) {
            m = fmax(m, F.cfl_part[2 * i]);
            s = s + F.cfl_part[2 * i + 1];
        }
        cmax = block_max<1024>(sh, m);
        q3s = block_sum<1024>(sh2, s);
    }
    if (t == 0) {
        StepScalars r;
        r.err = 0;
        if (n == 1) {
            double cflm = cmax, dt = dt_in;
            r.q3sum = q3s;
            if (!F.idtv) {
                if (ns == 0) {
                    cflm = cflm * dt;
                } else if (ns > 0) {
                    dt = F.cflmax / cflm;
                    if (dt > F.dtmax) dt = F.dtmax;
                }
                if (dt < 1.0e-8) r.err = FRIBO_E_DT_TOO_SMALL;
            } else {
                cflm = cflm * dt;
                if (cflm > F.cfllim) r.err = FRIBO_E_CFL_TOO_LARGE;
            }
            // an error set before the step (a batched instance whose markers
            // left the domain) stops it like the host check does
            if (s_err) r.err = s_err;
            r.dt = dt;
            r.cflm = cflm;
            r.beta = dt / F.ren * 0.5;
        } else {
            r.dt = s_dt;
            r.cflm = s_cflm;
            r.q3sum = 0.0;
            r.beta = s_beta;
            if (s_err) r.err = s_err;
        }
        double al, ga, ro;
        if (F.nsst == 3) {
            al = F.alm[n];
            ga = F.gam[n];
            ro = F.rom[n];
        } else {
            if (ns == 0) {
                ga = 1.0;
                ro = 0.0;
            } else {
                ga = 1.0 + 0.5 * r.dt / dt_o;
                ro = -0.5 * r.dt / dt_o;
            }
            al = ga + ro;
        }
        r.al = al;
        r.ga = ga;
        r.ro = ro;
        r.betadx = r.beta * al;
        out = r;
    }
    lds_barrier();
    return out;
}

// LDS slot of row r in sys_prepare: one pad slot per 32 rows, so that the
// lanes of the chunk elimination, each reading row cs[p] + i of its own
// chunk, fall on different banks (unpadded, 32-row chunks put all 64 lanes
// on one bank)
__device__ inline int sp_slot(int r) { return r + (r >> 5); }

// Grid coefficients (am, ac, ap) of system K.
__device__ inline void sys_coefs(const FluidDev& F, const KSys& K, const double*& am, const double*& ac,
                                 const double*& ap)
{
    if (K.src == 0) { am = F.am3sk; ac = F.ac3sk; ap = F.ap3sk; }
    else if (K.src == 1) { am = F.am2j; ac = F.ac2j; ap = F.ap2j; }
    else { am = F.am3j; ac = F.ac3j; ap = F.ap3j; }
}

// Line-invariant factors of one partitioned Helmholtz system for this
// substep (solve_fluid.F90:77-87,126-143,185-200,260-278: rows normalised by
// 1/(1 - betadx*ac)).  Row coefficients go to LDS; lane p of the first wave
// then eliminates chunk p (forward multipliers, left-coupling vector and the
// weights that turn forward values into first-row values), and the wave
// builds the separator system and its PCR multipliers.
__device__ void sys_prepare(const FluidDev& F, const KSys& K, double betadx, double* lds)
{
    const int t = threadIdx.x;
    const int n = K.n;
    const int ns = sp_slot(n - 1) + 1;
    double* la = lds;
    double* lc = lds + ns;
    double* lS = lds + 2 * ns;
    const double *am, *ac, *ap;
    sys_coefs(F, K, am, ac, ap);
    // rows t and t + 1024 per pass (workgroups of 1024 threads), both rows'
    // loads issued before the first row's stores, which the compiler would
    // otherwise keep ahead of the second row's loads (the coefficient arrays
    // and K's factor arrays may alias as far as it knows)
    for (int r0 = t; r0 < n; r0 += 2048) {
        double amr[2], acr[2], apr[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int idx = min(r0 + 1024 * u, n - 1) + K.kofs;
            amr[u] = am[idx];
            acr[u] = ac[idx];
            apr[u] = ap[idx];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int r = r0 + 1024 * u;
            if (r < n) {
                const double b = 1.0 / (1.0 - acr[u] * betadx);
                const double amv = -amr[u] * betadx * b;
                const double apv = -apr[u] * betadx * b;
                const double av = (r == 0) ? 0.0 : amv, cv = (r == n - 1) ? 0.0 : apv;
                const int q = sp_slot(r);
                la[q] = av;
                lc[q] = cv;
                lS[q] = b;
                if (r == 0) K.bc[0] = amv;
                if (r == n - 1) K.bc[1] = apv;
                if (!K.inter) {
                    K.S[r] = b;
                    K.a[r] = av;
                    K.c[r] = cv;
                }
            }
        }
    }
    lds_barrier();
    if (t >= 64) return;
    const int p = t;
    double alpha = 0.0, beta = 1.0, gamma = 0.0, vs = 0.0, ws = 0.0, ce = 0.0;
    if (p < K.nch) {
        const int s = K.cs[p], e = K.cs[p + 1] - 1;
        double cpp = 0.0, vpp = 1.0, gg = 1.0, wl = 0.0, glast = 1.0, vsum = 0.0;
#pragma unroll 4
        for (int i = s; i <= e - 1; ++i) {
            const double a = la[sp_slot(i)], c = lc[sp_slot(i)];
            const double iv = recip(1.0 - a * cpp);
            const double cpi = c * iv;
            const double vpi = (-a * vpp) * iv;
            const int idx = K.inter ? (i - s) * 64 + p : i;
            K.inv[idx] = iv;
            K.cp[idx] = cpi;
            K.vp[idx] = vpi;
            K.g[idx] = gg;
            if (K.inter) {
                K.S[idx] = lS[sp_slot(i)];
                K.a[idx] = a;
                K.c[idx] = c;
            }
            vsum = vsum + gg * vpi;
            if (i == e - 1) {
                wl = (-c) * iv;
                glast = gg;
            }
            gg = -cpi * gg;
            cpp = cpi;
            vpp = vpi;
        }
        const int se = sp_slot(e);
        if (K.inter) {
            const int idx = (e - s) * 64 + p;
            K.S[idx] = lS[se];
            K.a[idx] = la[se];
            K.c[idx] = lc[se];
        }
        K.wl[p] = wl;
        vs = vsum;
        ws = glast * wl;
        const double ae = la[se];
        ce = lc[se];
        alpha = ae * vpp;
        beta = 1.0 + ae * wl;
    }
    const double vs_n = __shfl_down(vs, 1, 64);
    const double ws_n = __shfl_down(ws, 1, 64);
    if (p + 1 < K.nch) {
        beta = beta + ce * vs_n;
        gamma = ce * ws_n;
    }
    int li = 0;
    for (int h = 1; h < 64; h <<= 1, ++li) {
        double amn = __shfl_up(alpha, h, 64), bm = __shfl_up(beta, h, 64), gm = __shfl_up(gamma, h, 64);
        double apn = __shfl_down(alpha, h, 64), bp = __shfl_down(beta, h, 64), gp = __shfl_down(gamma, h, 64);
        if (p - h < 0) { amn = 0.0; bm = 1.0; gm = 0.0; }
        if (p + h >= 64) { apn = 0.0; bp = 1.0; gp = 0.0; }
        const double m1 = alpha * recip(bm), m2 = gamma * recip(bp);
        K.k1[li * 64 + p] = m1;
        K.k2[li * 64 + p] = m2;
        const double na = -m1 * amn, ng = -m2 * gp;
        beta = beta - m1 * gm - m2 * apn;
        alpha = na;
        gamma = ng;
    }
    K.ib[p] = 1.0 / beta;
}

size_t sysprep_lds_bytes(const FluidDev& F)
{
    int n = F.ks[0].n;
    for (int i = 0; i < 2; ++i) {
        if (F.ks[i].n > n) n = F.ks[i].n;
        if (F.js[i].n > n) n = F.js[i].n;
    }
    return (size_t)3 * (n + (n >> 5) + 1) * sizeof(double);
}

// Pressure partials k_upd leaves: one per tile (the mirror form's tiles
// cover the top half with its mirror image, half as many).
__device__ inline int npr_count(const FluidDev& F)
{
    return F.ysh > 0 ? fribo_npr_part(F.n2m / 2, F.n3m) : F.npr_part;
}

// The first PR_PRE pressure partials of thread t (entries t, t+1024, ...,
// clamped to the last), loaded ahead so that they travel with the kernel's
// other loads (8192 partials: the 2048x1536 grid has 6400, 3328 in the
// mirror form).
#define PR_PRE 8
__device__ inline void pr_load(const FluidDev& F, double* v)
{
    const int t = threadIdx.x, n = npr_count(F);
#pragma unroll
    for (int u = 0; u < PR_PRE; ++u) v[u] = F.pr_part[min(t + 1024 * u, n - 1)];
}

// Pressure mean of the previous substep (pr_submean, updvp.F90:125-141) from
// the npr_part partials of k_upd: thread t adds entries t, t+1024, ... in
// turn (pr_partial: the first PR_PRE from pr_load, then eight loads in
// flight at a time), then block_sum.  Fixed order.
__device__ inline double pr_partial(const FluidDev& F, const double* v0)
{
    const int t = threadIdx.x, n = npr_count(F);
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < PR_PRE; ++u)
        if (t + 1024 * u < n) s = s + v0[u];
    for (int b = t + 1024 * PR_PRE; b < n; b += 8 * 1024) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = F.pr_part[min(b + u * 1024, n - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (b + u * 1024 < n) s = s + v[u];
    }
    return s;
}
__device__ inline double pr_mean(const FluidDev& F, double* sh, const double* v0)
{
    return block_sum<1024>(sh, pr_partial(F, v0)) / (F.rext2 * F.alx3);
}

// Three block_sum<1024> reductions in one pass (each value summed in the
// same order as block_sum, so the results are the same bits).
__device__ inline void block_sum3(double* sh, double& a, double& b, double& c)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a = a + __shfl_xor(a, o, 64);
        b = b + __shfl_xor(b, o, 64);
        c = c + __shfl_xor(c, o, 64);
    }
    lds_barrier();
    if ((t & 63) == 0) {
        sh[t >> 6] = a;
        sh[16 + (t >> 6)] = b;
        sh[32 + (t >> 6)] = c;
    }
    lds_barrier();
    a = sh[0];
    b = sh[16];
    c = sh[32];
#pragma unroll
    for (int w = 1; w < 16; ++w) {
        a = a + sh[w];
        b = b + sh[16 + w];
        c = c + sh[32 + w];
    }
}

// pr_submean and boucqt (iobcs.F90:81-108) by one workgroup of 1024
// threads.  The mean is kept pending for the interior (S->pend, subtracted on
// the next read) and accumulated for the entries no kernel reads (row j=n2
// and the ghost planes, S->goff, subtracted when the state is read back).
// Thread t owns rows j = 1+t, 1+t+1024, ...; the loads of two rows are issued
// before their stores.
__device__ void mean_block(const FluidDev& F, double* sh)
{
    DevScalars* S = F.S;
    const int t = threadIdx.x, n2m = F.n2m;
    double prv[PR_PRE];
    pr_load(F, prv);
    const double pravg = pr_mean(F, sh, prv);
    for (int j0 = 1 + t; j0 <= n2m; j0 += 2048) {
        double a[2][8];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int j = min(j0 + 1024 * u, n2m);
            a[u][0] = F.qb2s[j];
            a[u][1] = F.dqb2s[j];
            a[u][2] = F.qb3s[j];
            a[u][3] = F.dqb3s[j];
            a[u][4] = F.qb2n[j];
            a[u][5] = F.dqb2n[j];
            a[u][6] = F.qb3n[j];
            a[u][7] = F.dqb3n[j];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int j = j0 + 1024 * u;
            if (j <= n2m) {
                F.qb2s[j] = a[u][0] + a[u][1];
                F.qb3s[j] = a[u][2] + a[u][3];
                F.qb2n[j] = a[u][4] + a[u][5];
                F.qb3n[j] = a[u][6] + a[u][7];
            }
        }
    }
    if (t == 0) {
        S->pend = pravg;
        S->goff = S->goff + pravg;
    }
}

// Row values read by boucqt/boucdq for rows j0 and j0 + 1024 (clamped to n2m).
struct BcRows {
    double qn2[2], qn3[2], q2e[2], q3e[2], o2[2], o3[2], g[2];
    double qs2[2], qs3[2], dqs2[2], dqs3[2], dqn2[2], dqn3[2];
};

__device__ inline void bc_load(const FluidDev& F, int mean, int j0, BcRows& b)
{
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int j = min(j0 + 1024 * u, F.n2m);
        b.qn2[u] = F.qb2n[j];
        b.qn3[u] = F.qb3n[j];
        b.q2e[u] = F.q2[fidx(F, j, F.n3m)];
        b.q3e[u] = F.q3[fidx(F, j, F.n3m)];
        b.o2[u] = F.dq2x2o[j];
        b.o3[u] = F.dq3x2o[j];
        b.g[u] = F.g2rm[j];
        if (mean) {
            b.qs2[u] = F.qb2s[j];
            b.qs3[u] = F.qb3s[j];
            b.dqs2[u] = F.dqb2s[j];
            b.dqs3[u] = F.dqb3s[j];
            b.dqn2[u] = F.dqb2n[j];
            b.dqn3[u] = F.dqb3n[j];
        }
    }
}

// Workgroup 0 of k_bc: with mean, pr_submean/boucqt of the previous substep
// (as mean_block), then boucdq (iobcs.F90:22-50) and the global mass balance
// at the outflow (iobcs.F90:54-74).  Thread t owns rows j = 1+t, 1+t+1024,
// ...: every row value is loaded once, and the boundary values updated by
// boucqt feed boucdq from registers.  The rows of the first pass (pre) and
// the first pressure partials (prv) are loaded at kernel entry, before the
// step scalars are known, so that a single memory round trip serves the
// common case n2m <= 2048; the outflow increments of those rows stay in
// registers for the mass correction.
__device__ void bc_rows(const FluidDev& F, const StepScalars& r, int mean, double* sh, const BcRows& pre,
                        const double* prv)
{
    DevScalars* S = F.S;
    const int t = threadIdx.x, n2m = F.n2m;
    double psum = mean ? pr_partial(F, prv) : 0.0;
    // the sum is formed here: sunk to its use after the pass, it would keep
    // the preloaded partials live through the pass and spill
    asm volatile("" : "+v"(psum));
    const double dt = r.dt, ga = r.ga, ro = r.ro;
    const double udx3c_n3 = F.udx3c[F.n3];
    double area = 0.0, qout = 0.0;
    double keep[2] = {0.0, 0.0};
    // one pass over rows j0 and j0 + 1024 from the values in b
    auto pass = [&](const BcRows& b, int j0, bool first) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int j = j0 + 1024 * u;
            if (j <= n2m) {
                double qn2 = b.qn2[u], qn3 = b.qn3[u];
                if (mean) {
                    F.qb2s[j] = b.qs2[u] + b.dqs2[u];
                    F.qb3s[j] = b.qs3[u] + b.dqs3[u];
                    qn2 = qn2 + b.dqn2[u];
                    qn3 = qn3 + b.dqn3[u];
                    F.qb2n[j] = qn2;
                    F.qb3n[j] = qn3;
                }
                F.dqb2s[j] = 0.0;
                F.dqb3s[j] = 0.0;
                const double dq2x2 = (qn2 - b.q2e[u]) * 2.0 * udx3c_n3;
                const double d2 = -dt * (ga * dq2x2 + ro * b.o2[u]) * F.cou;
                const double dq3x2 = (qn3 - b.q3e[u]) * udx3c_n3;
                const double d3 = -dt * (ga * dq3x2 + ro * b.o3[u]) * F.cou;
                F.dqb2n[j] = d2;
                F.dqb3n[j] = d3;
                F.dq2x2o[j] = dq2x2;
                F.dq3x2o[j] = dq3x2;
                if (first) keep[u] = d3;
                const double darea = b.g[u] / F.dx2;
                area = area + darea;
                qout = qout + d3 * darea;
            }
        }
    };
    pass(pre, 1 + t, true);
    for (int j0 = 1 + t + 2048; j0 <= n2m; j0 += 2048) {
        BcRows b;
        bc_load(F, mean, j0, b);
        pass(b, j0, false);
    }
    // the pressure mean, the outflow area and the outflow flux in one reduction
    block_sum3(sh, psum, area, qout);
    const double tarea = area, tqout = qout;
    if (t == 0) {
        F.dqb2s[F.n2] = 0.0;
        F.dqb3s[F.n2] = 0.0;
        if (mean) {
            const double pravg = psum / (F.rext2 * F.alx3);
            S->pend = pravg;
            S->goff = S->goff + pravg;
        }
    }
    const double qinf = 0.0 * tarea;   // dqb3s == 0 (steady inflow)
    const double cor = (qinf - tqout) / tarea;
    int i = 0;
    for (int j = 1 + t; j <= n2m; j += 1024, ++i) F.dqb3n[j] = ((i < 2) ? keep[i] : F.dqb3n[j]) + cor;
    if (t == 0) {
        S->qinf = qinf;
        S->qout = tqout;
        S->area = tarea;
    }
}

// Workgroup 0: the previous substep's pressure mean and boucqt (mean != 0),
// finalize CFL/dt (substep 1), RK coefficients, boucdq.
// Workgroups 1..4: factors of the four Helmholtz systems of this substep.
__device__ __forceinline__ void bc_body(const FluidDev& F, const VBlock& vb, int n, int mean)
{
    extern __shared__ double dyn[];
    __shared__ double sh[1024];
    __shared__ double sh2[1024];
    DevScalars* S = F.S;
    const int t = threadIdx.x;
    const int ns = S->in_ns;
    const double dt_in = S->in_dt, dt_o_in = S->in_dto;
    const double s_dt = S->dt, s_cflm = S->cflm, s_beta = S->beta;
    const int s_err = S->err;
    // workgroup 0: its row values and pressure partials, in flight with the above
    BcRows pre;
    double prv[PR_PRE] = {};
    if (vb.x == 0) {
        bc_load(F, mean, 1 + t, pre);
        if (mean) pr_load(F, prv);
    }
    const StepScalars r = step_scalars(F, n, ns, dt_in, dt_o_in, s_dt, s_cflm, s_beta, s_err, sh, sh2);
    if (r.err) {
        if (vb.x == 0 && t == 0 && n == 1) {
            S->cflm = r.cflm;
            S->dt = r.dt;
            S->err = r.err;
        }
        return;
    }
    if (vb.x > 0) {
        const int sys = vb.x - 1;
        const KSys& K = (sys < 2) ? F.ks[sys] : F.js[sys - 2];
        sys_prepare(F, K, r.betadx, dyn);
        return;
    }
    if (t == 0) {
        if (n == 1) {
            S->q3avg = r.q3sum / (F.rext2 * F.alx3);
            S->cflm = r.cflm;
            S->dt = r.dt;
            S->dt_o = dt_o_in;
            S->beta = r.beta;
        }
        S->al = r.al;
        S->ga = r.ga;
        S->ro = r.ro;
        S->alre = r.al / F.ren;
        S->betadx = r.betadx;
        S->ugkks = 1.0 / (F.g3rm[1] * F.g3rc[1]) * F.dx3q * 2.0 * r.betadx;
        S->ugkkn = 1.0 / (F.g3rm[F.n3m] * F.g3rc[F.n3]) * F.dx3q * 2.0 * r.betadx;
    }
    bc_rows(F, r, mean, sh, pre, prv);
    // q2 wall rows: q2(1,:) = q2(n2,:) = 0 on the ghost planes (helsolve.F90:32-33)
    if (t < 6) {
        int k = (t < 3) ? t - 2 : F.n3m + 1 + (t - 3);
        F.q2[fidx(F, 1, k)] = 0.0;
        F.q2[fidx(F, F.n2, k)] = 0.0;
    }
}
__global__ __launch_bounds__(1024) void k_bc(const FluidDev F, int n, int mean)
{
    bc_body(F, vb_hw(), n, mean);
}
__global__ __launch_bounds__(1024) void kb_bc(const BGrid<FluidDev> G, int n, int mean)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    bc_body(*Fp, vb, n, mean);
}

void launch_bc(const FluidDev& F, int n, bool mean, hipStream_t st)
{
    hipLaunchKernelGGL(k_bc, dim3(5), dim3(1024), sysprep_lds_bytes(F), st, F, n, mean ? 1 : 0);
}

void launch_bc_b(BatchSet& B, int n, bool mean, hipStream_t st)
{
    size_t lds = 0;
    for (const auto& F : B.hF) lds = std::max(lds, sysprep_lds_bytes(F));
    blaunch(B, BK_BC, bdims(B, [](const FluidDev&) { return dim3(5); }), kb_bc, dim3(1024), lds, st, n,
            mean ? 1 : 0);
}

// ---------------------------------------------------------------------------
// Marching layout shared by k_rhs, k_div and k_upd: one 64-lane wave owns
// columns j = 62*bx + lane (lanes 1..62 produce output, lanes 0 and 63 are
// the j-1 / j+1 halo) and marches MK planes k = MK*by+1 .. MK*by+MK (MKR for
// k_rhs), keeping a register window of the planes it needs; j-neighbours come
// from the adjacent lanes, so every field value is read from HBM once.
#define MK FRIBO_MK
#define MKR FRIBO_MK_RHS
#define WJ FRIBO_WJ
// Waves (tiles) per workgroup: adjacent tiles in j share their boundary
// cache lines, which then come from the one L2 of the workgroup's XCD; the
// measured gain holds for k_div only (k_upd runs faster with one wave per
// workgroup, k_rhs is indifferent).
#ifndef SW_RHS
#define SW_RHS 1
#endif
#ifndef FRIBO_RHS_PF2
#define FRIBO_RHS_PF2 0
#endif
#ifndef SW_DIV
#define SW_DIV 8
#endif
#ifndef SW_UPD
#define SW_UPD 1
#endif
static_assert(MK <= 64 && MKR <= 64, "one plane of a tile per lane");

// Single-instance launches of the marching kernels: the hardware hands
// workgroups to the 8 XCDs round-robin by linear id, so the launch grid is
// padded to a multiple of 8 tile rows and workgroup id i takes tile
// (i % 8) * (n / 8) + i / 8 in row-major order.  Each XCD then works on a
// contiguous band of planes with all its columns: tiles adjacent in j (which
// share the cache lines at their edges) and most tiles adjacent in k (which
// share their halo planes) meet in one L2 instead of being fetched by two.
// Returns false for the padding workgroups.
__device__ inline bool vb_xcd_band(VBlock& vb, int gy_real)
{
    const int n = vb.gx * vb.gy;
    const int id = vb.x + vb.gx * vb.y;
    const int t = (id & 7) * (n >> 3) + (id >> 3);
    vb.x = t % vb.gx;
    vb.y = t / vb.gx;
    vb.gy = gy_real;
    return vb.y < gy_real;
}

static inline dim3 xcd_pad(dim3 g)
{
    return dim3(g.x, (g.y + 7) / 8 * 8);
}

__device__ inline double up1(double v) { return __shfl_up(v, 1, 64); }
__device__ inline double dn1(double v) { return __shfl_down(v, 1, 64); }

// Fused HDNL + INVTR.  Cell (j,k): q2 point (rc(j), zm(k)) for 2<=j<=n2m and
// q3 point (rm(j), zz(k)) for 2<=k<=n3m; operation order as in the reference
// loops (HDNL.F90:35-220, INVTR.F90:45-277).
__device__ __forceinline__ void rhs_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x & 63;
    const int ntj = (F.n2m + WJ - 1) / WJ, bx = (int)vb.x * SW_RHS + (int)(threadIdx.x >> 6);
    if (bx >= ntj) return;
    const int j = WJ * bx + lane;
    const int k0 = MKR * vb.y + 1;
    const int n2m = F.n2m, n3m = F.n3m;
    if (k0 > n3m) return;
    const int k1 = min(k0 + MKR - 1, n3m);
    const bool inj = (j >= 1 && j <= F.n2);
    const bool out = (lane >= 1 && lane <= WJ && j <= n2m);
    const double dt = S->dt, ga = S->ga, ro = S->ro, al = S->al, alre = S->alre, pend = S->pend;
    const double* __restrict__ q2 = F.q2;
    const double* __restrict__ q3 = F.q3;
    const double* __restrict__ pr = F.pr;
    double* __restrict__ ru2 = F.ru2;
    double* __restrict__ ru3 = F.ru3;
    double* __restrict__ rhs2 = F.rhs2;
    double* __restrict__ rhs3 = F.rhs3;
    const size_t P = F.P;
    size_t c = inj ? fidx(F, j, k0) : 0;
#if FRIBO_RHS_PF2
    // window: planes k-1, k, k+1 of q2/q3; the loads run two planes ahead
    // (q2/q3 of plane k+3, pr/ru2/ru3 of plane k+2 are issued in plane k)
    const bool two = k0 + 1 <= k1;
    double q2m = inj ? q2[c - P] : 0.0, q2c = inj ? q2[c] : 0.0;
    double q3m = inj ? q3[c - P] : 0.0, q3c = inj ? q3[c] : 0.0;
    double prm = inj ? pr[c - P] - pend : 0.0;
    double q2n = inj ? q2[c + P] : 0.0, q3n = inj ? q3[c + P] : 0.0;
    double prn = inj ? pr[c] - pend : 0.0;
    double ru2n = out ? ru2[c] : 0.0, ru3n = out ? ru3[c] : 0.0;
    double q2nn = (inj && two) ? q2[c + 2 * P] : 0.0, q3nn = (inj && two) ? q3[c + 2 * P] : 0.0;
    double prnn = (inj && two) ? pr[c + P] - pend : 0.0;
    double ru2nn = (out && two) ? ru2[c + P] : 0.0, ru3nn = (out && two) ? ru3[c + P] : 0.0;
#else
    // window: planes k-1, k; plane k+1 (and ru2/ru3/pr of plane k) are
    // loaded one iteration ahead so two planes of loads are in flight
    double q2m = inj ? q2[c - P] : 0.0, q2c = inj ? q2[c] : 0.0;
    double q3m = inj ? q3[c - P] : 0.0, q3c = inj ? q3[c] : 0.0;
    double prm = inj ? pr[c - P] - pend : 0.0;
    double q2n = inj ? q2[c + P] : 0.0, q3n = inj ? q3[c + P] : 0.0;
    double prn = inj ? pr[c] - pend : 0.0;
    double ru2n = out ? ru2[c] : 0.0, ru3n = out ? ru3[c] : 0.0;
#endif
    // per-column coefficients
    const int jj = out ? j : 1;
    const double udx2c = F.udx2c[jj], udx2m = F.udx2m[jj];
    const double am2 = F.am2j[jj], ac2 = F.ac2j[jj], ap2 = F.ap2j[jj];
    const double am3 = F.am3j[jj], ac3 = F.ac3j[jj], ap3 = F.ap3j[jj];
    const double qb2s = F.qb2s[jj], qb2n = F.qb2n[jj];
    const double qb3s = F.qb3s[jj], qb3sm = F.qb3s[jj > 1 ? jj - 1 : 1];
    const double qb3n = F.qb3n[jj], qb3nm = F.qb3n[jj > 1 ? jj - 1 : 1];
    // per-plane coefficients of the tile: lane l holds plane k0 + l
    const int kl = min(k0 + lane, n3m);
    const double l_udx3m = F.udx3m[kl], l_udx3c = F.udx3c[kl];
    const double l_ap3sk = F.ap3sk[kl], l_ac3sk = F.ac3sk[kl], l_am3sk = F.am3sk[kl];
    const double l_ap3ck = F.ap3ck[kl], l_ac3ck = F.ac3ck[kl], l_am3ck = F.am3ck[kl];
    for (int k = k0; k <= k1; ++k, c += P) {
        const double q2p = q2n, q3p = q3n, prc = prn, ru2c = ru2n, ru3c = ru3n;
#if FRIBO_RHS_PF2
        q2n = q2nn;
        q3n = q3nn;
        prn = prnn;
        ru2n = ru2nn;
        ru3n = ru3nn;
        if (k + 1 < k1) {
            q2nn = inj ? q2[c + 3 * P] : 0.0;
            q3nn = inj ? q3[c + 3 * P] : 0.0;
            prnn = inj ? pr[c + 2 * P] - pend : 0.0;
            ru2nn = out ? ru2[c + 2 * P] : 0.0;
            ru3nn = out ? ru3[c + 2 * P] : 0.0;
        }
#else
        if (k < k1) {
            q2n = inj ? q2[c + 2 * P] : 0.0;
            q3n = inj ? q3[c + 2 * P] : 0.0;
            prn = inj ? pr[c + P] - pend : 0.0;
            ru2n = out ? ru2[c + P] : 0.0;
            ru3n = out ? ru3[c + P] : 0.0;
        }
#endif
        const int kk = k - k0;
        const double udx3m_k = lane_bcast(l_udx3m, kk), udx3c_k = lane_bcast(l_udx3c, kk);
        const double ap3sk_k = lane_bcast(l_ap3sk, kk), ac3sk_k = lane_bcast(l_ac3sk, kk);
        const double am3sk_k = lane_bcast(l_am3sk, kk);
        const double ap3ck_k = lane_bcast(l_ap3ck, kk), ac3ck_k = lane_bcast(l_ac3ck, kk);
        const double am3ck_k = lane_bcast(l_am3ck, kk);
        // neighbours in j
        const double q2E = dn1(q2c), q2W = up1(q2c), q2Em = dn1(q2m), q2Ep = dn1(q2p);
        const double q3W = up1(q3c), q3Wp = up1(q3p), q3E = dn1(q3c);
        const double prW = up1(prc);
        (void)q2Ep;
        if (out) {
            // ---- q2: hdnl2 + invtr2 (j >= 2)
            if (j >= 2) {
                double a = q2E + q2c, b = q2W + q2c;
                double h22 = (a * a - b * b) * udx2c * 0.25;
                double dphv = -h22;
                double h23;
                if (k >= 2 && k <= n3m - 1) {
                    h23 = ((q3p + q3Wp) * (q2p + q2c) - (q3c + q3W) * (q2c + q2m)) * udx3m_k * 0.25;
                } else if (k == n3m) {
                    h23 = (qb2n * (qb3n + qb3nm) - (q3c + q3W) * (q2c + q2m) * 0.5) * udx3m_k * 0.5;
                } else {
                    h23 = ((q3p + q3Wp) * (q2c + q2p) * 0.5 - qb2s * (qb3s + qb3sm)) * udx3m_k * 0.5;
                }
                dphv = dphv - h23;
                double d22q2 = q2E * ap2 + q2c * ac2 + q2W * am2;
                double d33q2, dpx22;
                if (k >= 2 && k <= n3m - 1) {
                    d33q2 = q2p * ap3sk_k + q2c * ac3sk_k + q2m * am3sk_k;
                    double udx2 = al * udx2c;
                    dpx22 = (prc - prW) * udx2;
                } else if (k == 1) {
                    double codea = F.dx3q / (F.g3rc[2] * F.g3rm[1]);
                    double codeb = F.dx3q / (F.g3rc[1] * F.g3rm[1]);
                    d33q2 = ((q2p - q2c) * codea - (q2c - qb2s) * codeb * 2.0);
                    dpx22 = (prc - prW) * udx2c * al;
                } else {
                    double codea = F.dx3q / (F.g3rc[k + 1] * F.g3rm[k]);
                    double codeb = F.dx3q / (F.g3rc[k] * F.g3rm[k]);
                    d33q2 = ((qb2n - q2c) * codea * 2.0 - (q2c - q2m) * codeb);
                    dpx22 = (prc - prW) * udx2c * al;
                }
                double dcq2 = d22q2 + d33q2;
                rhs2[c] = (ga * dphv + ro * ru2c + alre * dcq2 - dpx22) * dt;
                ru2[c] = dphv;
            }
            // ---- q3: hdnl3 + invtr3 (k >= 2)
            if (k >= 2) {
                double qcapv = 0.0;   // qcap(:,n3m) is never written (stays 0 from initia)
                if (k <= n3m - 1) {
                    const double q3pp = (j == n2m) ? q3c : q3E;   // jpv clamps at n2m
                    const double q3mm = (j == 1) ? q3c : q3W;     // jmv clamps at 1
                    double h32 = (((q2E + q2Em) * (q3pp + q3c)) - ((q2c + q2m) * (q3c + q3mm))) * udx2m * 0.25;
                    double a = q3p + q3c, b = q3c + q3m;
                    double h33 = (a * a - b * b) * udx3c_k * 0.25;
                    qcapv = -(h32 + h33);
                }
                const double udx3 = al * udx3c_k;
                double dq32;
                if (j == 1) dq32 = q3E * ap3 + q3c * ac3 + F.q3lw * am3;
                else if (j == n2m) dq32 = F.q3uw * ap3 + q3c * ac3 + q3W * am3;
                else dq32 = q3E * ap3 + q3c * ac3 + q3W * am3;
                double dq33 = q3p * ap3ck_k + q3c * ac3ck_k + q3m * am3ck_k;
                double dcq3 = dq32 + dq33;
                double dpx33 = (prc - prm) * udx3;
                rhs3[c] = (ga * qcapv + ro * ru3c + alre * dcq3 - dpx33) * dt;
                ru3[c] = qcapv;
            }
        }
        q2m = q2c;
        q2c = q2p;
        q3m = q3c;
        q3c = q3p;
        prm = prc;
    }
}
__global__ __launch_bounds__(64 * SW_RHS) void k_rhs(const FluidDev F)
{
    VBlock vb = vb_hw();
    if (vb_xcd_band(vb, (F.n3m + MKR - 1) / MKR)) rhs_body(F, vb);
}
__global__ __launch_bounds__(64 * SW_RHS) void kb_rhs(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block_xcd(G, vb);
    if (Fp) rhs_body(*Fp, vb);
}

static dim3 grid_rhs(const FluidDev& F)
{
    const int ntj = (F.n2m + WJ - 1) / WJ;
    return dim3((ntj + SW_RHS - 1) / SW_RHS, (F.n3m + MKR - 1) / MKR);
}

void launch_rhs(const FluidDev& F, hipStream_t st)
{
    hipLaunchKernelGGL(k_rhs, xcd_pad(grid_rhs(F)), dim3(64 * SW_RHS), 0, st, F);
}

void launch_rhs_b(BatchSet& B, hipStream_t st)
{
    blaunch_xcd(B, BK_RHS, bdims(B, grid_rhs), kb_rhs, dim3(64 * SW_RHS), 0, st);
}

// ---------------------------------------------------------------------------
// divergence.F90:26-39
// MIRROR (the mirror form of the transform along y, n2m = 2h): the wave
// owns tile bx of the top half (rows up to h) and the mirrored tile of the
// bottom half, lane l on row n2m + 1 - j (descending, so the j+1 neighbour
// is in lane l-1), and stores the pair butterfly of the two divergences,
// dph(j) = d(j) + d(n2m+1-j), dph(n2m+1-j) = d(j) - d(n2m+1-j), instead of
// the divergences themselves.
template <bool MIRROR>
__device__ __forceinline__ void div_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x & 63;
    const int jtop = MIRROR ? F.n2m / 2 : F.n2m;   // last output row of the tiles
    const int ntj = (jtop + WJ - 1) / WJ, bx = (int)vb.x * SW_DIV + (int)(threadIdx.x >> 6);
    if (bx >= ntj) return;
    const int j = WJ * bx + lane;
    const int k0 = MK * vb.y + 1;
    if (k0 > F.n3m) return;
    const int k1 = min(k0 + MK - 1, F.n3m);
    const bool inj = (j >= 1 && j <= F.n2);
    const bool out = (lane >= 1 && lane <= WJ && j <= jtop);
    const double usdtal = 1.0 / (S->dt * S->al);
    const size_t P = F.P;
    size_t c = inj ? fidx(F, j, k0) : 0;
    const double udx2m = F.udx2m[out ? j : 1];
    const double* __restrict__ q2 = F.q2;
    const double* __restrict__ q3 = F.q3;
    double* __restrict__ dph = F.dph;
    double q3c = inj ? q3[c] : 0.0;
    double q2n = inj ? q2[c] : 0.0, q3n = inj ? q3[c + P] : 0.0;
    // the mirrored row (MIRROR only)
    const int jm = F.n2m + 1 - j;
    const bool injm = MIRROR && jm >= 1 && jm <= F.n2;
    size_t cm = injm ? fidx(F, jm, k0) : 0;
    const double udx2mm = MIRROR ? F.udx2m[out ? jm : 1] : 0.0;
    double q3cm = injm ? q3[cm] : 0.0;
    double q2nm = injm ? q2[cm] : 0.0, q3nm = injm ? q3[cm + P] : 0.0;
    const double l_udx3m = F.udx3m[min(k0 + lane, F.n3m)];   // lane l: plane k0 + l
    for (int k = k0; k <= k1; ++k, c += P, cm += P) {
        const double q2c = q2n, q3p = q3n;
        const double q2cm = q2nm, q3pm = q3nm;
        if (k < k1) {
            q2n = inj ? q2[c + P] : 0.0;
            q3n = inj ? q3[c + 2 * P] : 0.0;
            if (MIRROR) {
                q2nm = injm ? q2[cm + P] : 0.0;
                q3nm = injm ? q3[cm + 2 * P] : 0.0;
            }
        }
        const double q2E = dn1(q2c);
        const double q2Em = MIRROR ? up1(q2cm) : 0.0;
        const double udx3m_k = lane_bcast(l_udx3m, k - k0);
        if (out) {
            double dq = (q2E - q2c) * udx2m + (q3p - q3c) * udx3m_k;
            if (MIRROR) {
                const double dqm = (q2Em - q2cm) * udx2mm + (q3pm - q3cm) * udx3m_k;
                const double vt = dq * usdtal, vm = dqm * usdtal;
                dph[c] = vt + vm;
                dph[cm] = vt - vm;
            } else {
                dph[c] = dq * usdtal;
            }
        }
        q3c = q3p;
        q3cm = q3pm;
    }
}
template <bool MIRROR>
__global__ __launch_bounds__(64 * SW_DIV) void k_div(const FluidDev F)
{
    VBlock vb = vb_hw();
    if (vb_xcd_band(vb, (F.n3m + MK - 1) / MK)) div_body<MIRROR>(F, vb);
}
__global__ __launch_bounds__(64 * SW_DIV) void kb_div(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block_xcd(G, vb);
    if (!Fp) return;
    if (Fp->ysh > 0) div_body<true>(*Fp, vb);
    else div_body<false>(*Fp, vb);
}

static dim3 grid_div(const FluidDev& F)
{
    const int ntj = ((F.ysh > 0 ? F.n2m / 2 : F.n2m) + WJ - 1) / WJ;
    return dim3((ntj + SW_DIV - 1) / SW_DIV, (F.n3m + MK - 1) / MK);
}

// The single-instance step passes F.ysh = 0 unless the mirror form runs
// (transform along y, FP64); batches clear it for instances without it.
void launch_div(const FluidDev& F, hipStream_t st)
{
    if (F.ysh > 0) hipLaunchKernelGGL(k_div<true>, xcd_pad(grid_div(F)), dim3(64 * SW_DIV), 0, st, F);
    else hipLaunchKernelGGL(k_div<false>, xcd_pad(grid_div(F)), dim3(64 * SW_DIV), 0, st, F);
}

void launch_div_b(BatchSet& B, hipStream_t st)
{
    blaunch_xcd(B, BK_DIV, bdims(B, grid_div), kb_div, dim3(64 * SW_DIV), 0, st);
}

// ---------------------------------------------------------------------------
// updvp.F90:27-56 (velocity correction) + 83-102 (prcalc) + partial sums of
// the pressure mean (updvp.F90:127-133).  The stored pressure is "raw": the
// previous mean (S->pend) is subtracted on read.
// MIRROR (the mirror form of the transform along y, n2m = 2h): dph holds the
// backward transform's result before its pair butterfly, C; the wave owns
// tile bx of the top half (rows up to h) and the mirrored tile of the bottom
// half, lane l on row n2m + 1 - j (descending: its j-1 neighbour is in lane
// l+1, its j+1 neighbour in lane l-1), loads C on both rows and takes
// dph(j) = C(j) + C(n2m+1-j), dph(n2m+1-j) = C(j) - C(n2m+1-j) (for a halo
// row j past h the roles swap).
template <bool MIRROR>
__device__ __forceinline__ void upd_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x & 63;
    const int h = F.n2m / 2;
    const int jtop = MIRROR ? h : F.n2m;   // last output row of the tiles
    const int ntj = (jtop + WJ - 1) / WJ, bx = (int)vb.x * SW_UPD + (int)(threadIdx.x >> 6);
    if (bx >= ntj) return;
    const int j = WJ * bx + lane;
    const int k0 = MK * vb.y + 1;
    double part = 0.0;
    if (k0 <= F.n3m) {
        const int k1 = min(k0 + MK - 1, F.n3m);
        const bool inj = (j >= 1 && j <= F.n2);
        const bool out = (lane >= 1 && lane <= WJ && j <= jtop);
        const double al = S->al, dt = S->dt, pend = S->pend, be = al * S->beta;
        const size_t P = F.P;
        const int jj = out ? j : 1;
        const double usujm = al * dt * F.udx2c[jj];
        const double apj = F.apphj[jj], acj = F.acphj[jj], amj = F.amphj[jj];
        const double dy = F.rc[jj + 1] - F.rc[jj];
        const bool jlo = (j == 1), jhi = (j == F.n2m);
        size_t c = inj ? fidx(F, j, k0) : 0;
        // the mirrored row (MIRROR only)
        const int jm = F.n2m + 1 - j;
        const bool injm = MIRROR && jm >= 1 && jm <= F.n2;
        const int jjm = (MIRROR && out) ? jm : 1;
        const double usujm_m = MIRROR ? al * dt * F.udx2c[jjm] : 0.0;
        const double apj_m = MIRROR ? F.apphj[jjm] : 0.0, acj_m = MIRROR ? F.acphj[jjm] : 0.0;
        const double amj_m = MIRROR ? F.amphj[jjm] : 0.0;
        const double dy_m = MIRROR ? F.rc[jjm + 1] - F.rc[jjm] : 0.0;
        const bool jlo_m = (jm == 1), jhi_m = (jm == F.n2m);
        size_t cm = injm ? fidx(F, jm, k0) : 0;
        const bool lower = j <= h;   // row j in the top half: dph(j) = a + b
        const double* __restrict__ dph = F.dph;
        double* __restrict__ q2 = F.q2;
        double* __restrict__ q3 = F.q3;
        double* __restrict__ pr = F.pr;
        const int n3m = F.n3m;
        // gauge constant of the transform along y (0 along z): the backward
        // transform's partials, summed in tile order
        double doff = 0.0;
        if (F.taxis == 1) {
            double g = 0.0;
            for (int q = 0; q < F.ngpart; ++q) g = g + F.gpart[q];
            doff = g * F.zginv;
        }
        if (vb.x == 0 && vb.y == 0 && lane == 0) F.S->dphoff = doff;
        // dph of row j (t) and of the mirrored row (m) at plane offset o
        auto ld = [&](size_t o, bool ok, double& t, double& m) {
            if (!MIRROR) {
                t = (inj && ok) ? dph[c + o] - doff : 0.0;
                return;
            }
            const double a = (inj && ok) ? dph[c + o] : 0.0, b = (injm && ok) ? dph[cm + o] : 0.0;
            t = (lower ? a + b : b - a) - doff;
            m = (lower ? a - b : a + b) - doff;
        };
        double dm, dc, dn, dm_m = 0.0, dc_m = 0.0, dn_m = 0.0;
        ld(-P, k0 > 1, dm, dm_m);
        ld(0, true, dc, dc_m);
        // plane k+1 of dph and plane k of q2/q3/pr are loaded one iteration ahead
        ld(P, k0 < n3m, dn, dn_m);
        double q2n = out ? q2[c] : 0.0, q3n = out ? q3[c] : 0.0, prn = out ? pr[c] : 0.0;
        double q2n_m = 0.0, q3n_m = 0.0, prn_m = 0.0;
        if (MIRROR && out) {
            q2n_m = q2[cm];
            q3n_m = q3[cm];
            prn_m = pr[cm];
        }
        // per-plane coefficients of the tile: lane l holds plane k0 + l
        const int kl = min(k0 + lane, n3m);
        const double l_udx3c = F.udx3c[kl], l_apphk = F.apphk[kl], l_acphk = F.acphk[kl];
        const double l_amphk = F.amphk[kl], l_dz = F.zz[kl + 1] - F.zz[kl];
        for (int k = k0; k <= k1; ++k, c += P, cm += P) {
            const int kk = k - k0;
            const double udx3c_k = lane_bcast(l_udx3c, kk), apphk_k = lane_bcast(l_apphk, kk);
            const double acphk_k = lane_bcast(l_acphk, kk), amphk_k = lane_bcast(l_amphk, kk);
            const double dz_k = lane_bcast(l_dz, kk);
            const double dp = dn, q2v = q2n, q3v = q3n, prv0 = prn;
            const double dp_m = dn_m, q2v_m = q2n_m, q3v_m = q3n_m, prv0_m = prn_m;
            if (k < k1) {
                ld(2 * P, k + 1 < n3m, dn, dn_m);
                q2n = out ? q2[c + P] : 0.0;
                q3n = out ? q3[c + P] : 0.0;
                prn = out ? pr[c + P] : 0.0;
                if (MIRROR && out) {
                    q2n_m = q2[cm + P];
                    q3n_m = q3[cm + P];
                    prn_m = pr[cm + P];
                }
            }
            const double dW = up1(dc), dE = dn1(dc);
            const double dW_m = MIRROR ? dn1(dc_m) : 0.0, dE_m = MIRROR ? up1(dc_m) : 0.0;
            if (out) {
                const double usukm = al * dt * udx3c_k;
                {
                    const double dJm = jlo ? dc : dW;     // dph(jmv(j))
                    const double dJp = jhi ? dc : dE;     // dph(jpv(j))
                    const double dKm = (k == 1) ? dc : dm; // dph(kmv(k))
                    const double dKp = (k == n3m) ? dc : dp;
                    q2[c] = q2v - (dc - dJm) * usujm;
                    if (k >= 2) q3[c] = q3v - (dc - dm) * usukm;
                    double lap = (dJp * apj + dc * acj + dJm * amj) +
                                 (dKp * apphk_k + dc * acphk_k + dKm * amphk_k);
                    double prv = (prv0 - pend) + dc - be * lap;
                    pr[c] = prv;
                    part = part + prv * dz_k * dy;
                }
                if (MIRROR) {
                    const double dJm = jlo_m ? dc_m : dW_m;
                    const double dJp = jhi_m ? dc_m : dE_m;
                    const double dKm = (k == 1) ? dc_m : dm_m;
                    const double dKp = (k == n3m) ? dc_m : dp_m;
                    q2[cm] = q2v_m - (dc_m - dJm) * usujm_m;
                    if (k >= 2) q3[cm] = q3v_m - (dc_m - dm_m) * usukm;
                    double lap = (dJp * apj_m + dc_m * acj_m + dJm * amj_m) +
                                 (dKp * apphk_k + dc_m * acphk_k + dKm * amphk_k);
                    double prv = (prv0_m - pend) + dc_m - be * lap;
                    pr[cm] = prv;
                    part = part + prv * dz_k * dy_m;
                }
            }
            dm = dc;
            dc = dp;
            dm_m = dc_m;
            dc_m = dp_m;
        }
    }
    // wave reduction (fixed order) -> one partial per workgroup
    for (int off = 32; off > 0; off >>= 1) part = part + __shfl_xor(part, off, 64);
    if (lane == 0) F.pr_part[vb.y * ntj + bx] = part;
}
template <bool MIRROR>
__global__ __launch_bounds__(64 * SW_UPD) void k_upd(const FluidDev F)
{
    VBlock vb = vb_hw();
    if (vb_xcd_band(vb, (F.n3m + MK - 1) / MK)) upd_body<MIRROR>(F, vb);
}
__global__ __launch_bounds__(64 * SW_UPD) void kb_upd(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block_xcd(G, vb);
    if (!Fp) return;
    if (Fp->ysh > 0) upd_body<true>(*Fp, vb);
    else upd_body<false>(*Fp, vb);
}

static dim3 grid_upd(const FluidDev& F)
{
    const int ntj = ((F.ysh > 0 ? F.n2m / 2 : F.n2m) + WJ - 1) / WJ;
    return dim3((ntj + SW_UPD - 1) / SW_UPD, (F.n3m + MK - 1) / MK);
}

void launch_upd(const FluidDev& F, hipStream_t st)
{
    if (F.ysh > 0) hipLaunchKernelGGL(k_upd<true>, xcd_pad(grid_upd(F)), dim3(64 * SW_UPD), 0, st, F);
    else hipLaunchKernelGGL(k_upd<false>, xcd_pad(grid_upd(F)), dim3(64 * SW_UPD), 0, st, F);
}

void launch_upd_b(BatchSet& B, hipStream_t st)
{
    blaunch_xcd(B, BK_UPD, bdims(B, grid_upd), kb_upd, dim3(64 * SW_UPD), 0, st);
}

// ---------------------------------------------------------------------------
// pr_submean (updvp.F90:125-141) and boucqt (iobcs.F90:81-108) as a kernel
// of its own (after the last substep of a step, and after every substep of
// the substep-by-substep entry point); within a step, k_bc of the next
// substep does the same work first (bc_rows).
__device__ __forceinline__ void mean_body(const FluidDev& F, const VBlock& vb)
{
    __shared__ double sh[16];
    if (F.S->err) return;
    mean_block(F, sh);
}
__global__ __launch_bounds__(1024) void k_mean(const FluidDev F)
{
    mean_body(F, vb_hw());
}
__global__ __launch_bounds__(1024) void kb_mean(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    mean_body(*Fp, vb);
}

void launch_mean(const FluidDev& F, hipStream_t st)
{
    hipLaunchKernelGGL(k_mean, dim3(1), dim3(1024), 0, st, F);
}

void launch_mean_b(BatchSet& B, hipStream_t st)
{
    blaunch(B, BK_MEAN, bdims(B, [](const FluidDev&) { return dim3(1); }), kb_mean, dim3(1024), 0, st);
}

// ---------------------------------------------------------------------------
// ttime's per-step diagnostics (ttime.F90:29-69,136) on the device: one pass
// over q2/q3 gives vmaxv/vminv (vcheck.F90:3-73), divgck
// (divergence.F90:49-78), kinen (utils_proc.F90:474-515) and cfl
// (cfl.F90:19-38) as per-block partials; k_diag_final reduces them in a
// fixed order and adds getintvort (utils_proc.F90:582-629) on the plane kv
// nearest to z = 18.5 (0: outside the grid).  out[9] = {vmax2, vmax3, vmin2,
// vmin3, max|div|, ektot, vorticity integral, cflm, q3avg}.
#define DIAG_NB 256
__global__ __launch_bounds__(256) void k_diag(const FluidDev F, double* part)
{
    __shared__ double sh[256];
    const int t = threadIdx.x;
    const int per = (F.n3m + DIAG_NB - 1) / DIAG_NB;
    const int k0 = blockIdx.x * per + 1, k1 = min(k0 + per - 1, F.n3m);
    double v2 = -100.0, v3 = -100.0, m2 = 100.0, m3 = 100.0, qm = -100.0, ek = 0.0, cf = 1.0e-8, qs = 0.0;
    for (int k = k0; k <= k1; ++k) {
        const double u3 = F.udx3m[k], dz = F.zz[k + 1] - F.zz[k];
        for (int j = 1 + t; j <= F.n2m; j += 256) {
            const size_t c = fidx(F, j, k);
            const double a2 = F.q2[c], b2 = F.q2[c + 1], a3 = F.q3[c], b3 = F.q3[c + F.P];
            v2 = fmax(v2, fabs(a2));
            v3 = fmax(v3, fabs(a3));
            m2 = fmin(m2, a2);
            m3 = fmin(m3, a3);
            const double dq = (b2 - a2) * F.udx2m[j] + (b3 - a3) * u3;
            qm = fmax(fabs(dq), qm);
            const double dy = F.rc[j + 1] - F.rc[j];
            const double vol = dz * dy;
            const double vccv = (a2 + b2) * 0.5, vccw = (a3 + b3) * 0.5;
            ek = ek + (vccv * vccv + vccw * vccw) * vol;
            cf = fmax(cf, fabs((a2 + b2) * 0.5 * F.udx2m[j]) + fabs((a3 + b3) * 0.5 * u3));
            qs = qs + a3 * dz * dy;
        }
    }
    double vals[8] = {v2, v3, m2, m3, qm, ek, cf, qs};
    for (int q = 0; q < 8; ++q) {
        sh[t] = vals[q];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) {
                const double a = sh[t], b = sh[t + h];
                sh[t] = (q == 2 || q == 3) ? fmin(a, b) : ((q == 5 || q == 7) ? a + b : fmax(a, b));
            }
            __syncthreads();
        }
        if (t == 0) part[q * DIAG_NB + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_diag_final(const FluidDev F, const double* part, int kv, double* out)
{
    __shared__ double sh[256];
    const int t = threadIdx.x;
    double r[8];
    for (int q = 0; q < 8; ++q) {
        sh[t] = part[q * DIAG_NB + t];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) {
                const double a = sh[t], b = sh[t + h];
                sh[t] = (q == 2 || q == 3) ? fmin(a, b) : ((q == 5 || q == 7) ? a + b : fmax(a, b));
            }
            __syncthreads();
        }
        r[q] = sh[0];
        __syncthreads();
    }
    double vi = 0.0;
    if (kv > 0) {
        const double dz = F.zz[kv + 1] - F.zz[kv];
        for (int j = 1 + t; j <= F.n2m; j += 256) {
            const double dyj = F.rc[j + 1] - F.rc[j];
            const double dwdy = (F.q3[fidx(F, j + 1, kv)] - F.q3[fidx(F, j, kv)]) / dyj;
            const double dvdz = (F.q2[fidx(F, j, kv + 1)] - F.q2[fidx(F, j, kv)]) / dz;
            vi = vi + (dwdy - dvdz) * dyj;
        }
    }
    sh[t] = vi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) sh[t] = sh[t] + sh[t + h];
        __syncthreads();
    }
    if (t == 0) {
        out[0] = r[0];
        out[1] = r[1];
        out[2] = r[2];
        out[3] = r[3];
        out[4] = r[4];
        out[5] = 0.5 * r[5] / (F.rext2 * F.alx3);
        out[6] = sh[0];
        out[7] = r[6];
        out[8] = r[7] / (F.rext2 * F.alx3);
    }
}

void launch_diag(const FluidDev& F, double* part, int kv, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_diag, dim3(DIAG_NB), dim3(256), 0, st, F, part);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, F, (const double*)part, kv, out);
}

// ---------------------------------------------------------------------------
// Residual of the pressure Poisson system for the mixed-precision solve:
// r = rhs - (A_y + A_z) x over the n2m x n3m cells, with the tridiagonal
// operators the transforms invert (phini.F90:121-142: rows amphj/acphj/apphj
// along y, amphk/acphk/apphk along z, the couplings past the first and last
// rows dropped).  One thread per column j marches RES_MK planes with x of
// planes k-1, k, k+1 in registers; r may alias rhs.
#define RES_MK 16
__global__ __launch_bounds__(256) void k_pres_residual(const FluidDev F, const double* __restrict__ x,
                                                       const double* rhs, double* r)
{
    if (F.S->err) return;
    const int j = blockIdx.x * 256 + threadIdx.x + 1;
    if (j > F.n2m) return;
    const int k0 = blockIdx.y * RES_MK + 1, k1 = min(k0 + RES_MK - 1, F.n3m);
    const double am = j > 1 ? F.amphj[j] : 0.0, ac = F.acphj[j], ap = j < F.n2m ? F.apphj[j] : 0.0;
    const bool jl = j > 1, jr = j < F.n2m;
    double xm = k0 > 1 ? x[fidx(F, j, k0 - 1)] : 0.0;
    double xc = x[fidx(F, j, k0)];
    for (int k = k0; k <= k1; ++k) {
        const size_t c = fidx(F, j, k);
        const double xp = k < F.n3m ? x[c + F.P] : 0.0;
        const double xw = jl ? x[c - 1] : 0.0, xe = jr ? x[c + 1] : 0.0;
        const double bk = k > 1 ? F.amphk[k] : 0.0, ck = F.acphk[k], tk = k < F.n3m ? F.apphk[k] : 0.0;
        const double ly = am * xw + ac * xc + ap * xe;
        const double lz = bk * xm + ck * xc + tk * xp;
        r[c] = rhs[c] - (ly + lz);
        xm = xc;
        xc = xp;
    }
}

void launch_pres_residual(const FluidDev& F, const double* x, const double* rhs, double* r, hipStream_t st)
{
    hipLaunchKernelGGL(k_pres_residual, dim3((F.n2m + 255) / 256, (F.n3m + RES_MK - 1) / RES_MK), dim3(256), 0, st,
                       F, x, rhs, r);
}
