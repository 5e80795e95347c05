// Moving-least-squares immersed-boundary kernels (gfx950).
//
// Substep 1 of every step (fluid_tsch.F90:116-118, forcing_MLS.F90:59-96):
//   k_mls_idx   lgrdistribution.F90:39-75  nearest Eulerian nodes of every
//               marker (Fortran MINLOC: first index of the minimum), one wave
//               per (marker, search), exact FP64 |x - p| comparisons
//   k_mls_sup   MLS_SD_ind (MLSint.F90:4-104), MLS_SF + weight_function +
//               compute_AB + GEsolve (MLSint.F90:115-166,284-486,
//               utils_math.f90:637-725) and the momentum coefficient cl
//               (forcing_MLS.F90:72-81); one thread per (marker, direction)
//   k_mls_order orders the 9 contributions of every marker by (cell, marker)
//               (counting sort over the supports' bounding box of cells, or a
//               stable LDS radix sort on the cell; bitonic LDS blocks plus
//               global merge passes for very large marker sets); equal cells
//               form contiguous segments
// After the step, on request (fluid_tsch.F90:167-177):
//   k_loads      loads_transfer_c (loads_transfer.F90:4-162): probes at
//                ddmin*kprobe along the normal, MLS_SF_d shape functions and
//                derivatives (MLSint.F90:173-258,342-402,492-585), getpbup
//                gathers, pressure and friction stress per marker
//   k_checkbcs   checkbcs (ckbcs.F90:5-86)
// Every substep (interp_spread, forcing_MLS.F90:135-182):
//   k_mls_interp  u_m = sum phi q*f, f = (V - u_m)/dt, contribution cl*phi*f
//                 (sixteen lanes per marker and direction, one per support)
//   k_mls_spread  one thread per touched cell sums its contributions in
//                 marker order (the order the reference accumulates fey/fez
//                 in), zeroes plane k = n3m (the single-rank ghost update of
//                 mpi_routines.F90:324-339) and adds f*dt to rhs2/rhs3
//                 (helsolve.F90:20-49).  No atomics, bit-reproducible.
#include "fribo_batch.h"

#include <algorithm>

#define LMCC(n, c) F.lmcc[(size_t)((n) - 1) + (size_t)((c) - 1) * F.nlm]
#define LMVEL(n, c) F.lmvel[(size_t)((n) - 1) + (size_t)((c) - 1) * F.nlm]
#define LMIND(n, d, c) F.lmind[(size_t)((n) - 1) + (size_t)((d) - 1) * F.nlm + (size_t)((c) - 1) * 2 * F.nlm]
#define EULINS(n, d, c, m)                                                                   \
    F.eulins[(size_t)((n) - 1) + (size_t)((d) - 1) * F.nlm + (size_t)((c) - 1) * 2 * F.nlm + \
             (size_t)((m) - 1) * 4 * F.nlm]
#define PHI(n, d, m) F.phi[(size_t)((n) - 1) + (size_t)((d) - 1) * F.nlm + (size_t)((m) - 1) * 2 * F.nlm]
#define PHIPRED(n, d, m) F.phipred[(size_t)((n) - 1) + (size_t)((d) - 1) * F.nlm + (size_t)((m) - 1) * 2 * F.nlm]
#define CL(n, d) F.cl[(size_t)((n) - 1) + (size_t)((d) - 1) * F.nlm]

static const unsigned long long KEY_NONE = ~0ull;
#define ENT_BITS 20
#define ENT_MASK ((1ull << ENT_BITS) - 1)

// Fortran MINLOC(|x(1:len) - v|) over a 1-based, strictly increasing array,
// one wave.  |x(i) - v| computed in FP64 falls then rises along i (rounding is
// monotone), so the first minimum lies in the stretch that brackets v: the
// wave samples every step-th entry, keeps the stretch [s, s+step] with
// x(s) <= v (s = 1 if v < x(1)) and takes the first minimum there, with the
// same comparisons as a scan of the whole array.
__device__ int wave_minloc(const double* x, int len, double v)
{
    const int lane = threadIdx.x;
    const int step = (len + 63) / 64;
    const int si = 1 + lane * step;
    const bool le = si <= len && x[si] <= v;
    const unsigned long long bal = __ballot(le);
    const int s = bal ? 1 + (63 - __clzll(bal)) * step : 1;
    const int hi = min(len, s + step);
    double best = 0.0;
    int bi = 0x7fffffff;
    for (int i = s + lane; i <= hi; i += 64) {
        double a = fabs(x[i] - v);
        if (bi == 0x7fffffff || a < best) {
            best = a;
            bi = i;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        double ob = __shfl_xor(best, off, 64);
        int oi = __shfl_xor(bi, off, 64);
        bool take = (oi != 0x7fffffff) && (bi == 0x7fffffff || ob < best || (ob == best && oi < bi));
        if (take) {
            best = ob;
            bi = oi;
        }
    }
    return bi;
}

__device__ __forceinline__ void mls_idx_body(const FluidDev& F, const VBlock& vb)
{
    if (F.S->err) return;
    const int n = vb.x + 1, s = vb.y;
    const double y = LMCC(n, 1), z = LMCC(n, 2);
    int ind;
    if (s == 0) {
        ind = wave_minloc(F.rc, F.n2, y);
    } else if (s == 1) {
        if (z < F.zm[1] && z > 0.0) ind = 1;
        else if (z > F.zm[F.n3m] && z < F.alx3) ind = F.n3m;
        else ind = wave_minloc(F.zm, F.n3, z);
    } else if (s == 2) {
        if (y < F.rm[1] && y > 0.0) ind = 1;
        else if (y > F.rm[F.n2m] && y < F.rext2) ind = F.n2m;
        else ind = wave_minloc(F.rm, F.n2, y);
    } else {
        ind = wave_minloc(F.zz, F.n3, z);
    }
    if (threadIdx.x == 0) {
        int d = (s < 2) ? 1 : 2, c = (s & 1) + 1;
        LMIND(n, d, c) = ind;
    }
}
__global__ __launch_bounds__(64) void k_mls_idx(const FluidDev F)
{
    mls_idx_body(F, vb_hw());
}
__global__ __launch_bounds__(64) void kb_mls_idx(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    mls_idx_body(*Fp, vb);
}

// utils_math.f90:637-725, 3x3 scaled partial pivoting; returns info
__device__ int gesolve3(const double ai[3][3], const double bi[3], double x[3])
{
    double a[3][3], b[3], sc[3];
    for (int i = 0; i < 3; ++i) {
        b[i] = bi[i];
        for (int j = 0; j < 3; ++j) a[i][j] = ai[i][j];
    }
    for (int k = 0; k < 2; ++k) {
        for (int i = k; i < 3; ++i) {
            sc[i] = 0.0;
            for (int j = k; j < 3; ++j) sc[i] = fmax(sc[i], fabs(a[i][j]));
        }
        double pivot = fabs(a[k][k] / sc[k]);
        int l = k;
        for (int j = k + 1; j < 3; ++j)
            if (fabs(a[j][k] / sc[j]) > pivot) {
                pivot = fabs(a[j][k] / sc[j]);
                l = j;
            }
        if (pivot == 0.0) return 0;
        if (l != k) {
            for (int j = k; j < 3; ++j) {
                double st = a[k][j];
                a[k][j] = a[l][j];
                a[l][j] = st;
            }
            double st = b[k];
            b[k] = b[l];
            b[l] = st;
        }
        for (int i = k + 1; i < 3; ++i) {
            double c = a[i][k] / a[k][k];
            a[i][k] = 0.0;
            b[i] = b[i] - c * b[k];
            for (int j = k + 1; j < 3; ++j) a[i][j] = a[i][j] - c * a[k][j];
        }
    }
    x[2] = b[2] / a[2][2];
    for (int i = 1; i >= 0; --i) {
        double c = 0.0;
        for (int j = i + 1; j < 3; ++j) c = c + a[i][j] * x[j];
        x[i] = (b[i] - c) / a[i][i];
    }
    return 1;
}

// MLSint.F90:406-486
__device__ double weight1d(double dif, double sdd, int wfun, double radiusSD)
{
    const double wcon = 0.3;
    double sdds = radiusSD * sdd;
    double rcv = fabs(dif) / sdds;
    double rc2 = rcv * rcv, rc3 = rcv * rcv * rcv, rc4 = rcv * rcv * rcv * rcv;
    if (wfun == 1) {
        if (rcv > 1.0) return 0.0;
        if (rcv <= 0.5) return (2.0 / 3.0) - 4.0 * rc2 + 4.0 * rc3;
        return (4.0 / 3.0) - 4.0 * rcv + 4.0 * rc2 - (4.0 / 3.0) * rc3;
    } else if (wfun == 2) {
        double t = rcv / wcon;
        return (rcv <= 1.0) ? exp(-(t * t)) : 0.0;
    } else if (wfun == 3) {
        return (rcv <= 1.0) ? 1.0 - 6.0 * rc2 + 8.0 * rc3 - 3.0 * rc4 : 0.0;
    } else if (wfun == 4) {
        if (rcv > 1.0) return 0.0;
        if (rcv < 0.5) return 1.0 / 3.0 * (1.0 + sqrt(-3.0 * rc2 + 1.0));
        double om = 1.0 - rcv;
        return 1.0 / 6.0 * (5.0 - 3.0 * rcv - sqrt(-3.0 * (om * om) + 1.0));
    }
    return 0.0;
}

// Sixteen lanes per (marker, direction), lane m < 9 owning support m (MLS_SD_ind's
// order: j outer, k inner): each lane finds its support node, cell size,
// weight and monomials; the 3x3 moment matrix, the partition-of-unity sum and
// the two momentum sums are summed in support order from the group's lanes,
// so every value is the serial loop's; each lane of the group solves the same
// 3x3 system and keeps its own shape-function value.
__device__ __forceinline__ void mls_sup_body(const FluidDev& F, const VBlock& vb)
{
    DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x;
    const int g = vb.x * 4 + (lane >> 4), m = lane & 15, base = lane & 48;
    if (g >= 2 * F.nlm) return;   // whole group
    const int n = g % F.nlm + 1, d = g / F.nlm + 1;
    const bool sup = m < FRIBO_DIMSD;
    const double pos[2] = {LMCC(n, 1), LMCC(n, 2)};
    unsigned long long* keys = F.keys + (size_t)(d - 1) * F.maxkeys;
    const bool act = (F.zz[1] < pos[1]) && (pos[1] < F.zz[F.n3]);
    if (d == 1 && m == 0) F.active[n - 1] = act ? 1 : 0;
    if (!act) {                   // whole group
        if (sup) {
            PHI(n, d, m + 1) = 0.0;
            PHIPRED(n, d, m + 1) = 0.0;
            keys[(size_t)(n - 1) * FRIBO_DIMSD + m] = KEY_NONE;
        }
        if (m == 0) CL(n, d) = 0.0;
        return;
    }
    // MLS_SD_ind: support mm of the 3x3 block (lanes past the ninth repeat it)
    const int i1 = LMIND(n, d, 1), i2 = LMIND(n, d, 2);
    const double* ya = (d == 1) ? F.rc : F.rm;
    const double* za = (d == 1) ? F.zm : F.zz;
    const int mm = sup ? m : FRIBO_DIMSD - 1, j = mm / 3 + 1, k = mm % 3 + 1;
    const int jj = (i1 == 1) ? j + 1 : ((i1 == F.n2m) ? j - 1 : j);
    const int kk = (i2 == 1) ? k + 1 : ((i2 == F.n3m) ? k - 1 : k);
    const int ej = i1 + jj - 2, ek = i2 + kk - 2;
    const double sdc0 = ya[ej], sdd0 = ya[ej + 1] - ya[ej];
    const double sdc1 = za[ek], sdd1 = za[ek + 1] - za[ek];
    // MLS_SF
    const double pk0 = 1.0, pk1 = sdc0, pk2 = sdc1;
    const double w = weight1d(pos[0] - sdc0, sdd0, F.wfun, F.radiusSD) *
                     weight1d(pos[1] - sdc1, sdd1, F.wfun, F.radiusSD);
    const double t00 = w * (pk0 * pk0), t01 = w * (pk0 * pk1), t02 = w * (pk0 * pk2);
    const double t11 = w * (pk1 * pk1), t12 = w * (pk1 * pk2), t22 = w * (pk2 * pk2);
    double a00 = 0.0, a01 = 0.0, a02 = 0.0, a11 = 0.0, a12 = 0.0, a22 = 0.0;
#pragma unroll
    for (int i = 0; i < FRIBO_DIMSD; ++i) {
        a00 = a00 + __shfl(t00, base + i, 64);
        a01 = a01 + __shfl(t01, base + i, 64);
        a02 = a02 + __shfl(t02, base + i, 64);
        a11 = a11 + __shfl(t11, base + i, 64);
        a12 = a12 + __shfl(t12, base + i, 64);
        a22 = a22 + __shfl(t22, base + i, 64);
    }
    // A is symmetric entry by entry: w*(pa*pb) == w*(pb*pa)
    const double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    const double gps[3] = {1.0, pos[0], pos[1]};
    double ptA[3] = {0.0, 0.0, 0.0};
    gesolve3(A, gps, ptA);
    const double phl = ptA[0] * (pk0 * w) + ptA[1] * (pk1 * w) + ptA[2] * (pk2 * w);
    // forcing_MLS.F90:72-81 terms of this support
    const double tv = phl * (sdd0 * sdd1);
    const double tl = 1.0 / 2.0 * phl * (sdd0 + sdd1);
    double sumphi = 0.0, dvel = 0.0, lgr_h = 0.0;
#pragma unroll
    for (int i = 0; i < FRIBO_DIMSD; ++i) {
        sumphi = sumphi + __shfl(phl, base + i, 64);
        dvel = dvel + __shfl(tv, base + i, 64);
        lgr_h = lgr_h + __shfl(tl, base + i, 64);
    }
    if (m == 0) {
        if (fabs(1.0 - fabs(sumphi)) > 1.0e-10) atomicCAS(&S->err, 0, FRIBO_E_PARTITION_UNITY);
        CL(n, d) = F.lmA[n - 1] * lgr_h / dvel;
    }
    if (!sup) return;
    PHI(n, d, m + 1) = phl;
    PHIPRED(n, d, m + 1) = phl;
    EULINS(n, d, 1, m + 1) = ej;
    EULINS(n, d, 2, m + 1) = ek;
    // supports past the last node exist only for markers within half a cell
    // of the top wall / outflow, where the reference writes outside its
    // arrays; they are not spread
    const bool inb = ej >= 1 && ej <= F.n2 && ek >= 1 && ek <= F.n3;
    const unsigned long long cell = (unsigned long long)(ej - 1) + (unsigned long long)(ek - 1) * F.n2;
    keys[(size_t)(n - 1) * FRIBO_DIMSD + m] =
        inb ? ((cell << ENT_BITS) | (unsigned long long)((n - 1) * FRIBO_DIMSD + m)) : KEY_NONE;
}
__global__ __launch_bounds__(64) void k_mls_sup(const FluidDev F)
{
    mls_sup_body(F, vb_hw());
}
__global__ __launch_bounds__(64) void kb_mls_sup(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    mls_sup_body(*Fp, vb);
}

// Ordering of the (cell, entry) keys of each direction, ascending.
//
// Up to RADIX_MAX keys (the usual case: nine supports per marker) one
// workgroup per direction orders them (k_mls_order): a counting sort over the
// bounding box of the supports' cells when the box is small, else a stable
// LSD radix sort on the cell field (8-bit digits, as many passes as the grid's
// cell count needs).  For the radix sort: the keys are
// generated in entry order, so stability yields (cell, entry) order; an unused
// support (KEY_NONE) takes the largest digit value and lands at the end.  Each
// pass holds the keys in registers (thread-contiguous runs of 64-key rounds per
// wave), builds per-wave digit histograms, scans them digit-major and
// scatters into one LDS buffer; equal digits within a round are ranked with a
// peer mask built from eight ballots, and only the lowest lane of each peer
// group touches the wave's histogram counter.
//
// Larger key sets sort blocks of SORT_BLK keys in LDS (bitonic) and run the
// strides of the larger merge stages as global compare-exchange passes.
// Both orders are unique, so the result is deterministic.
#define SORT_BLK 2048
#define RADIX_MAX 16384
#define RADIX_WAVES 16
#define RADIX_R (RADIX_MAX / (64 * RADIX_WAVES))   // rounds of 64 keys per wave

// index of the first element of compare-exchange pair q at a power-of-two stride
__device__ inline int pair_lo(int q, int stride) { return ((q & ~(stride - 1)) << 1) | (q & (stride - 1)); }

// Radix path of k_mls_order: sk (RADIX_MAX keys), hist (256 x RADIX_WAVES)
// and wsum (RADIX_WAVES) are the workgroup's LDS.
__device__ void radix_sort_block(unsigned long long* __restrict__ gk, int n, int nbits, unsigned long long* sk,
                                 unsigned int* hist, unsigned int* wsum)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = ((n + 64 * RADIX_WAVES - 1) / (64 * RADIX_WAVES)) * 64;   // keys per wave
    const unsigned int none = (nbits >= 32) ? 0xffffffffu : ((1u << nbits) - 1u);
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int shift = 0; shift < nbits; shift += 8) {
        unsigned long long v[RADIX_R];
#pragma unroll
        for (int r = 0; r < RADIX_R; ++r) {
            const int e = w * per + r * 64 + lane;
            const bool ok = (r * 64 < per) && e < n;
            v[r] = ok ? (shift == 0 ? gk[e] : sk[e]) : KEY_NONE;
        }
        // digit of key r; 256 marks a lane without a key
        auto digit = [&](int r) -> unsigned int {
            const int e = w * per + r * 64 + lane;
            if (!((r * 64 < per) && e < n)) return 256u;
            const unsigned int cell = (v[r] == KEY_NONE) ? none : (unsigned int)(v[r] >> ENT_BITS);
            return (cell >> shift) & 255u;
        };
        // lanes of this round holding the same digit
        auto peers = [&](unsigned int dg) -> unsigned long long {
            unsigned long long m = __ballot(dg < 256u);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const unsigned long long bb = __ballot((dg >> b) & 1u);
                m &= ((dg >> b) & 1u) ? bb : ~bb;
            }
            return m;
        };
        for (int i = t; i < 256 * RADIX_WAVES; i += 1024) hist[i] = 0u;
        __syncthreads();
        // per-wave histogram: the lowest lane of each peer group adds the group size
#pragma unroll
        for (int r = 0; r < RADIX_R; ++r) {
            if (r * 64 >= per) break;
            const unsigned int dg = digit(r);
            const unsigned long long m = peers(dg);
            if (dg < 256u && (m & below) == 0ull)
                __hip_atomic_fetch_add(&hist[dg * RADIX_WAVES + w], (unsigned int)__popcll(m), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();
        // exclusive scan of the 4096 counters in [digit][wave] order: four per thread
        unsigned int c0 = hist[4 * t], c1 = hist[4 * t + 1], c2 = hist[4 * t + 2], c3 = hist[4 * t + 3];
        const unsigned int tot = c0 + c1 + c2 + c3;
        unsigned int inc = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        unsigned int base = 0u;
        for (int i = 0; i < w; ++i) base += wsum[i];
        const unsigned int ex = base + inc - tot;
        hist[4 * t] = ex;
        hist[4 * t + 1] = ex + c0;
        hist[4 * t + 2] = ex + c0 + c1;
        hist[4 * t + 3] = ex + c0 + c1 + c2;
        __syncthreads();
        // stable scatter: rounds in order, lanes in order within a round.  The
        // group leaders claim their ranges with returning LDS adds, issued for
        // all rounds back to back (LDS applies one wave's operations in
        // order, so a digit's claims follow the round order); the members
        // then take the leader's base by a shuffle.
        unsigned long long pm[RADIX_R];
        unsigned int pb[RADIX_R];
#pragma unroll
        for (int r = 0; r < RADIX_R; ++r) {
            pm[r] = 0ull;
            pb[r] = 0u;
            if (r * 64 < per) {
                const unsigned int dg = digit(r);
                pm[r] = peers(dg);
                if (dg < 256u && (pm[r] & below) == 0ull)
                    pb[r] = __hip_atomic_fetch_add(&hist[dg * RADIX_WAVES + w], (unsigned int)__popcll(pm[r]),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
#pragma unroll
        for (int r = 0; r < RADIX_R; ++r) {
            if (r * 64 >= per) break;
            const unsigned int dg = digit(r);
            const int lead = pm[r] ? __ffsll((long long)pm[r]) - 1 : lane;
            const unsigned int b = (unsigned int)__shfl((int)pb[r], lead, 64);
            if (dg < 256u) sk[b + (unsigned int)__popcll(pm[r] & below)] = v[r];
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += 1024) gk[i] = sk[i];
}

// Block-wide exclusive scan of one value per thread (1024 threads); returns
// the thread's prefix, *total gets the sum.  wsum: RADIX_WAVES words of LDS.
__device__ inline unsigned int block_exscan(unsigned int v, unsigned int* wsum, unsigned int* total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    __syncthreads();
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    unsigned int base = 0u, all = 0u;
    for (int i = 0; i < RADIX_WAVES; ++i) {
        if (i < w) base += wsum[i];
        all += wsum[i];
    }
    *total = all;
    return base + inc - v;
}

// Ordering of one direction's keys (one workgroup of 1024 threads).  The
// supports of a body span a small box of cells: when its W x H cells fit
// CELL_MAX, every key is counted into its box cell (LDS atomics), the counts
// are scanned into cell ranges, each key claims a slot of its range and
// records its entry there, and finally each key's place within its cell is
// its rank among the cell's entries -- the (cell, entry) order, whatever the
// order of the claims.  Larger boxes take the radix sort.
#define CELL_MAX 16384
__device__ __forceinline__ void mls_order_body(const FluidDev& F, const VBlock& vb, int nbits)
{
    __shared__ unsigned long long sk[RADIX_MAX];         // radix keys, or cnt[CELL_MAX] + ent[RADIX_MAX] (u32)
    __shared__ unsigned int hist[256 * RADIX_WAVES];
    __shared__ unsigned int wsum[RADIX_WAVES];
    __shared__ int box[4];
    if (F.S->err) return;
    const int d = vb.x, t = threadIdx.x, lane = t & 63;
    const int n = F.nlm * FRIBO_DIMSD;
    unsigned long long* gk = F.keys + (size_t)d * F.maxkeys;
    const unsigned int n2u = (unsigned int)F.n2;
    constexpr int KPT = RADIX_MAX / 1024;                // keys per thread
    unsigned long long key[KPT];
    int jmin = 0x7fffffff, jmax = -1, kmin = 0x7fffffff, kmax = -1;
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
        const int i = t + 1024 * u;
        key[u] = i < n ? gk[i] : KEY_NONE;
    }
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
        if (key[u] != KEY_NONE) {
            const unsigned int cell = (unsigned int)(key[u] >> ENT_BITS);
            const int j = (int)(cell % n2u), k = (int)(cell / n2u);
            jmin = min(jmin, j);
            jmax = max(jmax, j);
            kmin = min(kmin, k);
            kmax = max(kmax, k);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        jmin = min(jmin, __shfl_xor(jmin, o, 64));
        jmax = max(jmax, __shfl_xor(jmax, o, 64));
        kmin = min(kmin, __shfl_xor(kmin, o, 64));
        kmax = max(kmax, __shfl_xor(kmax, o, 64));
    }
    if (t < 4) box[t] = t & 1 ? -1 : 0x7fffffff;
    __syncthreads();
    if (lane == 0) {
        atomicMin(&box[0], jmin);
        atomicMax(&box[1], jmax);
        atomicMin(&box[2], kmin);
        atomicMax(&box[3], kmax);
    }
    __syncthreads();
    const int j0 = box[0], k0 = box[2];
    const int W = box[1] >= box[0] ? box[1] - box[0] + 1 : 0;
    const int H = box[3] >= box[2] ? box[3] - box[2] + 1 : 0;
    if ((long)W * H > CELL_MAX) {
        radix_sort_block(gk, n, nbits, sk, hist, wsum);
        return;
    }
    const int nb = W * H;
    unsigned int* cnt = reinterpret_cast<unsigned int*>(sk);
    unsigned int* ent = cnt + CELL_MAX;
    for (int i = t; i < nb; i += 1024) cnt[i] = 0u;
    __syncthreads();
    int lc[KPT];
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
        lc[u] = -1;
        if (key[u] != KEY_NONE) {
            const unsigned int cell = (unsigned int)(key[u] >> ENT_BITS);
            lc[u] = ((int)(cell % n2u) - j0) + ((int)(cell / n2u) - k0) * W;
            __hip_atomic_fetch_add(&cnt[lc[u]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
    // cell ranges: thread t scans cells [t*per, (t+1)*per)
    const int per = (nb + 1023) / 1024;
    unsigned int loc = 0u;
    for (int i = t * per; i < min(nb, (t + 1) * per); ++i) loc += cnt[i];
    unsigned int nvalid;
    unsigned int run = block_exscan(loc, wsum, &nvalid);
    for (int i = t * per; i < min(nb, (t + 1) * per); ++i) {
        const unsigned int c = cnt[i];
        cnt[i] = run;
        run += c;
    }
    __syncthreads();
    // claims: afterwards cnt[c] is the end of cell c's range
#pragma unroll
    for (int u = 0; u < KPT; ++u)
        if (lc[u] >= 0) {
            const unsigned int pos =
                __hip_atomic_fetch_add(&cnt[lc[u]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ent[pos] = (unsigned int)(key[u] & ENT_MASK);
        }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KPT; ++u)
        if (lc[u] >= 0) {
            const unsigned int e = cnt[lc[u]], s0 = lc[u] > 0 ? cnt[lc[u] - 1] : 0u;
            const unsigned int me = (unsigned int)(key[u] & ENT_MASK);
            // eight reads in flight per step (the range holds a few dozen entries)
            unsigned int r8[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
            unsigned int q = s0;
            for (; q + 8 <= e; q += 8) {
#pragma unroll
                for (int i = 0; i < 8; ++i) r8[i] += ent[q + i] < me ? 1u : 0u;
            }
            for (; q < e; ++q) r8[0] += ent[q] < me ? 1u : 0u;
            const unsigned int rank = ((r8[0] + r8[1]) + (r8[2] + r8[3])) + ((r8[4] + r8[5]) + (r8[6] + r8[7]));
            gk[s0 + rank] = key[u];
        }
    for (int i = (int)nvalid + t; i < n; i += 1024) gk[i] = KEY_NONE;
}
__global__ __launch_bounds__(1024) void k_mls_order(const FluidDev F, int nbits)
{
    mls_order_body(F, vb_hw(), nbits);
}
__global__ __launch_bounds__(1024) void kb_mls_order(const BGrid<FluidDev> G, int nbits)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    mls_order_body(*Fp, vb, nbits);
}

__global__ __launch_bounds__(1024) void k_mls_sort_local(const FluidDev F, int size0, int first)
{
    __shared__ unsigned long long sk[SORT_BLK];
    if (F.S->err) return;
    const int d = blockIdx.y, t = threadIdx.x;
    const int base = blockIdx.x * SORT_BLK;
    const int nreal = F.nlm * FRIBO_DIMSD;
    unsigned long long* gk = F.keys + (size_t)d * F.maxkeys;
    for (int i = t; i < SORT_BLK; i += 1024) {
        const int gi = base + i;
        sk[i] = (!first || gi < nreal) ? gk[gi] : KEY_NONE;
    }
    __syncthreads();
    // first pass: full sort of the block (sizes 2..SORT_BLK); later passes:
    // the strides < SORT_BLK of merge stage size0
    const int smin = first ? 2 : size0, smax = first ? SORT_BLK : size0;
    for (int size = smin; size <= smax; size <<= 1) {
        for (int stride = (size < SORT_BLK ? size : SORT_BLK) >> 1; stride > 0; stride >>= 1) {
            for (int q = t; q < SORT_BLK / 2; q += 1024) {
                const int i = pair_lo(q, stride);
                const int pj = i + stride;
                const bool up = (((base + i) & size) == 0);
                const unsigned long long x = sk[i], y = sk[pj];
                if ((x > y) == up) {
                    sk[i] = y;
                    sk[pj] = x;
                }
            }
            __syncthreads();
        }
    }
    for (int i = t; i < SORT_BLK; i += 1024) gk[base + i] = sk[i];
}

__global__ __launch_bounds__(256) void k_mls_sort_global(const FluidDev F, int size, int stride)
{
    if (F.S->err) return;
    const int d = blockIdx.y;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= F.maxkeys / 2) return;
    unsigned long long* gk = F.keys + (size_t)d * F.maxkeys;
    const int i = pair_lo(q, stride);
    const int pj = i + stride;
    const bool up = ((i & size) == 0);
    const unsigned long long x = gk[i], y = gk[pj];
    if ((x > y) == up) {
        gk[i] = y;
        gk[pj] = x;
    }
}

void launch_mls_setup(const FluidDev& F, hipStream_t st)
{
    if (F.nlm <= 0) return;
    hipLaunchKernelGGL(k_mls_idx, dim3(F.nlm, 4), dim3(64), 0, st, F);
    hipLaunchKernelGGL(k_mls_sup, dim3((2 * F.nlm + 3) / 4), dim3(64), 0, st, F);
    const unsigned long long ncell = (unsigned long long)F.n2 * (unsigned long long)F.n3;
    if (F.nlm * FRIBO_DIMSD <= RADIX_MAX && ncell < 0xffffffffull) {
        int nbits = 1;
        while (nbits < 32 && (1ull << nbits) <= ncell) ++nbits;   // 2^nbits > every cell + 1
        hipLaunchKernelGGL(k_mls_order, dim3(2), dim3(1024), 0, st, F, nbits);
        return;
    }
    const int nblk = F.maxkeys / SORT_BLK;
    hipLaunchKernelGGL(k_mls_sort_local, dim3(nblk, 2), dim3(1024), 0, st, F, 0, 1);
    for (int size = 2 * SORT_BLK; size <= F.maxkeys; size <<= 1) {
        for (int stride = size / 2; stride >= SORT_BLK; stride >>= 1)
            hipLaunchKernelGGL(k_mls_sort_global, dim3((F.maxkeys / 2 + 255) / 256, 2), dim3(256), 0, st, F, size,
                               stride);
        hipLaunchKernelGGL(k_mls_sort_local, dim3(nblk, 2), dim3(1024), 0, st, F, size, 0);
    }
}

// ---------------------------------------------------------------------------
// Sixteen lanes per (marker, direction), lane m < 9 owning support m: each
// lane gathers its own q + rhs value, the group's first lane sums the nine
// products in support order (the reference's order) from its neighbours, and
// every support lane writes its contribution.
__device__ __forceinline__ void mls_interp_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x;
    const int g = vb.x * 4 + (lane >> 4), m = lane & 15;   // group (n,d), support m
    const bool ing = g < 2 * F.nlm;
    const int n = (ing ? g % F.nlm : 0) + 1, d = (ing ? g / F.nlm : 0) + 1;
    const bool act = ing && F.active[n - 1];
    const bool sup = act && m < FRIBO_DIMSD;
    const double* q = (d == 1) ? F.q2 : F.q3;
    const double* r = (d == 1) ? F.rhs2 : F.rhs3;
    double prod = 0.0, ph = 0.0;
    if (sup) {
        const int j = EULINS(n, d, 1, m + 1), k = EULINS(n, d, 2, m + 1);
        // q2f is written for 2<=j<=n2m (INVTR.F90:145-149), q3f for
        // 2<=k<=n3m (INVTR.F90:273-277); elsewhere it keeps its zero start
        const bool def = (d == 1) ? (j >= 2 && j <= F.n2m && k >= 1 && k <= F.n3m)
                                  : (j >= 1 && j <= F.n2m && k >= 2 && k <= F.n3m);
        const size_t c = def ? fidx(F, j, k) : 0;
        const double qf = def ? q[c] + r[c] : 0.0;
        ph = PHI(n, d, m + 1);
        prod = ph * qf;
    }
    const int base = lane & ~15;
    double u = 0.0;
#pragma unroll
    for (int t = 0; t < FRIBO_DIMSD; ++t) u = u + __shfl(prod, base + t, 64);
    if (!sup) return;
    const double f = (LMVEL(n, d) - u) / S->dt;
    double* con = F.contrib + (size_t)(d - 1) * F.nlm * FRIBO_DIMSD;
    con[(size_t)(n - 1) * FRIBO_DIMSD + m] = CL(n, d) * ph * f;
}
__global__ __launch_bounds__(64) void k_mls_interp(const FluidDev F)
{
    mls_interp_body(F, vb_hw());
}
__global__ __launch_bounds__(64) void kb_mls_interp(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    mls_interp_body(*Fp, vb);
}

// One thread per sorted entry; the first entry of each cell (segment head)
// sums the cell's contributions in marker order.
__device__ __forceinline__ void mls_spread_body(const FluidDev& F, const VBlock& vb)
{
    const DevScalars* S = F.S;
    if (S->err) return;
    const int d = vb.y;
    const int i = vb.x * 64 + threadIdx.x;
    if (i >= F.nlm * FRIBO_DIMSD) return;
    const unsigned long long* gk = F.keys + (size_t)d * F.maxkeys;
    const unsigned long long key = gk[i];
    if (key == KEY_NONE) return;
    const unsigned long long cell = key >> ENT_BITS;
    if (i > 0 && (gk[i - 1] >> ENT_BITS) == cell) return;
    const double* con = F.contrib + (size_t)d * F.nlm * FRIBO_DIMSD;
    const int nreal = F.nlm * FRIBO_DIMSD;
    // the segment is read eight keys at a time (independent loads), summed in order
    double acc = 0.0;
    bool open = true;
    for (int e0 = i; open && e0 < nreal; e0 += 8) {
        unsigned long long kb[8];
        double cv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) kb[u] = (e0 + u < nreal) ? gk[e0 + u] : KEY_NONE;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool in = kb[u] != KEY_NONE && (kb[u] >> ENT_BITS) == cell;
            cv[u] = in ? con[kb[u] & ENT_MASK] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool in = open && kb[u] != KEY_NONE && (kb[u] >> ENT_BITS) == cell;
            if (in) acc = acc + cv[u];
            open = in;
        }
    }
    const unsigned cl32 = (unsigned)cell, n2u = (unsigned)F.n2;
    const int j = (int)(cl32 % n2u) + 1, k = (int)(cl32 / n2u) + 1;
    const double fval = (k == F.n3m) ? 0.0 : acc;
    F.segval[(size_t)d * F.maxkeys + i] = fval;
    if (k == F.n3m) return;
    if (d == 0) {
        if (j >= 2 && j <= F.n2m && k >= 1) {
            const size_t c = fidx(F, j, k);
            F.rhs2[c] = F.rhs2[c] + fval * S->dt;
        }
    } else {
        if (j >= 1 && j <= F.n2m && k >= 2) {
            const size_t c = fidx(F, j, k);
            F.rhs3[c] = F.rhs3[c] + fval * S->dt;
        }
    }
}
__global__ __launch_bounds__(64) void k_mls_spread(const FluidDev F)
{
    mls_spread_body(F, vb_hw());
}
__global__ __launch_bounds__(64) void kb_mls_spread(const BGrid<FluidDev> G)
{
    VBlock vb;
    const FluidDev* Fp = bg_block(G, vb);
    mls_spread_body(*Fp, vb);
}

void launch_mls_force(const FluidDev& F, hipStream_t st)
{
    if (F.nlm <= 0) return;
    hipLaunchKernelGGL(k_mls_interp, dim3((2 * F.nlm + 3) / 4), dim3(64), 0, st, F);
    hipLaunchKernelGGL(k_mls_spread, dim3((F.nlm * FRIBO_DIMSD + 63) / 64, 2), dim3(64), 0, st, F);
}

// Batched forms: instances without markers (or without the immersed
// boundary) get no blocks.
static bool has_markers(const FluidDev& F) { return F.ibact && F.nlm > 0; }

static int order_bits(const FluidDev& F)
{
    const unsigned long long ncell = (unsigned long long)F.n2 * (unsigned long long)F.n3;
    int nbits = 1;
    while (nbits < 32 && (1ull << nbits) <= ncell) ++nbits;   // 2^nbits > every cell + 1
    return nbits;
}

bool mls_batchable(const FluidDev& F)
{
    const unsigned long long ncell = (unsigned long long)F.n2 * (unsigned long long)F.n3;
    return !has_markers(F) || (F.nlm * FRIBO_DIMSD <= RADIX_MAX && ncell < 0xffffffffull);
}

void launch_mls_setup_b(BatchSet& B, hipStream_t st)
{
    int nbits = 1;
    for (const auto& F : B.hF)
        if (has_markers(F)) nbits = std::max(nbits, order_bits(F));
    blaunch(B, BK_MLS_IDX, bdims(B, [](const FluidDev& F) { return has_markers(F) ? dim3(F.nlm, 4) : dim3(0); }),
            kb_mls_idx, dim3(64), 0, st);
    blaunch(B, BK_MLS_SUP,
            bdims(B, [](const FluidDev& F) { return has_markers(F) ? dim3((2 * F.nlm + 3) / 4) : dim3(0); }),
            kb_mls_sup, dim3(64), 0, st);
    // a larger bit count than an instance's own only adds empty radix passes
    blaunch(B, BK_MLS_ORDER, bdims(B, [](const FluidDev& F) { return has_markers(F) ? dim3(2) : dim3(0); }),
            kb_mls_order, dim3(1024), 0, st, nbits);
}

void launch_mls_force_b(BatchSet& B, hipStream_t st)
{
    blaunch(B, BK_MLS_INTERP,
            bdims(B, [](const FluidDev& F) { return has_markers(F) ? dim3((2 * F.nlm + 3) / 4) : dim3(0); }),
            kb_mls_interp, dim3(64), 0, st);
    blaunch(B, BK_MLS_SPREAD, bdims(B, [](const FluidDev& F) {
                return has_markers(F) ? dim3((F.nlm * FRIBO_DIMSD + 63) / 64, 2) : dim3(0);
            }),
            kb_mls_spread, dim3(64), 0, st);
}

// ---------------------------------------------------------------------------
// Loads on the body (loads_transfer.F90:4-162) and the forcing check.

#define LMNOR(n, c) F.lmnor[(size_t)((n) - 1) + (size_t)((c) - 1) * F.nlm]
#define LMACC(n, c) F.lmacc[(size_t)((n) - 1) + (size_t)((c) - 1) * F.nlm]
#define LW(n, q) F.lwork[(size_t)((n) - 1) * FRIBO_LWORK + (q)]

// MLSint.F90:492-585 (weight_function_d), one direction: value and derivative
__device__ inline void weight1d_d(double difc, double sdd, int wfun, double radiusSD, double& wc, double& dw)
{
    const double wcon = 0.3, eps = 1.0e-20;
    const double sdds = radiusSD * sdd;
    const double drdc = (fabs(difc) <= eps) ? 0.0 : (difc / fabs(difc)) / sdds;
    const double rcv = fabs(difc) / sdds;
    const double rc2 = rcv * rcv, rc3 = rcv * rcv * rcv, rc4 = rcv * rcv * rcv * rcv;
    wc = 0.0;
    dw = 0.0;
    if (wfun == 1) {
        if (rcv > 1.0) {
        } else if (rcv <= 0.5) {
            wc = (2.0 / 3.0) - 4.0 * rc2 + 4.0 * rc3;
            dw = (-8.0 * rcv + 12.0 * rc2) * drdc;
        } else {
            wc = (4.0 / 3.0) - 4.0 * rcv + 4.0 * rc2 - (4.0 / 3.0) * rc3;
            dw = (-4.0 + 8.0 * rcv - 4.0 * rc2) * drdc;
        }
    } else if (wfun == 2) {
        if (rcv <= 1.0) {
            const double t = rcv / wcon;
            wc = exp(-(t * t));
            dw = (-1.0 / (wcon * wcon) * 2.0 * rcv * exp(-(t * t))) * drdc;
        }
    } else if (wfun == 3) {
        if (rcv <= 1.0) {
            wc = 1.0 - 6.0 * rc2 + 8.0 * rc3 - 3.0 * rc4;
            dw = -12.0 * rcv + 24.0 * rc2 - 12.0 * rc3;
        }
    } else if (wfun == 4) {
        if (rcv > 1.0) {
        } else if (rcv < 0.5) {
            wc = 1.0 / 3.0 * (1.0 + sqrt(-3.0 * rc2 + 1.0));
            dw = -rcv * pow(-3.0 * rc2 - 1.0, -0.5);
        } else {
            const double om = 1.0 - rcv;
            wc = 1.0 / 6.0 * (5.0 - 3.0 * rcv - sqrt(-3.0 * (om * om) + 1.0));
            dw = -(1.0 - rcv) / (2.0 * sqrt(1.0 - 3.0 * (om * om))) - 0.5;
        }
    }
}

// first minimum of |x(i) - v| over i = c-5..c+5 inside 1..n (probeind)
__device__ inline int window_minloc(const double* x, int n, int c, double v)
{
    int best = 0;
    double bv = 0.0;
    for (int i = c - 5; i <= c + 5; ++i) {
        if (i < 1 || i > n) continue;
        const double a = fabs(x[i] - v);
        if (best == 0 || a < bv) {
            bv = a;
            best = i;
        }
    }
    return best;
}

// Probe values: ucy/ucz/pc (loads_transfer.F90:210-267) need f(j-1,k) at j=1,
// which the reference's column-major storage resolves to f(n2,k-1).
__device__ inline double fat(const FluidDev& F, const double* f, int j, int k)
{
    if (j == 0) {
        j = F.n2;
        k = k - 1;
    }
    return f[fidx(F, j, k)];
}

__device__ inline double prat(const FluidDev& F, const DevScalars* S, int j, int k)
{
    if (j == 0) {
        j = F.n2;
        k = k - 1;
    }
    const bool edge = (j == F.n2) || k < 1 || k > F.n3m;
    return F.pr[fidx(F, j, k)] - (edge ? S->goff : S->pend);
}

// One thread per marker: probes, MLS_SF_d, getpbup and the pressure; the
// velocity-derivative combinations are left in lwork for the friction stress,
// which needs the tangent of the previous marker when this one has none.
__global__ __launch_bounds__(64) void k_loads(const FluidDev F, double hpr)
{
    DevScalars* S = F.S;
    if (S->err) return;
    const int n = blockIdx.x * 64 + threadIdx.x + 1;
    if (n > F.nlm) return;
    const double pos[2] = {LMCC(n, 1), LMCC(n, 2)};
    const double nor[2] = {LMNOR(n, 1), LMNOR(n, 2)};
    const double probe[2] = {pos[0] + nor[0] * hpr, pos[1] + nor[1] * hpr};
    F.press[n - 1] = 0.0;
    LW(n, 0) = 0.0;   // inside
    LW(n, 1) = 0.0;   // has a tangent
    if (!(F.zz[1] < pos[1] && pos[1] < F.zz[F.n3])) return;
    double u[3][7];
    for (int idir = 1; idir <= 3; ++idir) {
        const int mi1 = (idir == 3) ? LMIND(n, 2, 1) : LMIND(n, idir, 1);
        const int mi2 = (idir == 3) ? LMIND(n, 1, 2) : LMIND(n, idir, 2);
        const double* ya = (idir == 1) ? F.rc : F.rm;
        const double* za = (idir == 2) ? F.zz : F.zm;
        const int i1 = window_minloc(ya, F.n2, mi1, probe[0]);
        const int i2 = window_minloc(za, F.n3, mi2, probe[1]);
        // MLS_SD_ind
        double sdc[2][FRIBO_DIMSD], sdd[2][FRIBO_DIMSD];
        int eul[2][FRIBO_DIMSD];
        int m = 0;
        for (int j = 1; j <= 3; ++j)
            for (int k = 1; k <= 3; ++k) {
                const int jj = (i1 == 1) ? j + 1 : ((i1 == F.n2m) ? j - 1 : j);
                const int kk = (i2 == 1) ? k + 1 : ((i2 == F.n3m) ? k - 1 : k);
                const int ej = i1 + jj - 2, ek = i2 + kk - 2;
                sdc[0][m] = ya[ej];
                sdd[0][m] = ya[ej + 1] - ya[ej];
                sdc[1][m] = za[ek];
                sdd[1][m] = za[ek + 1] - za[ek];
                eul[0][m] = ej;
                eul[1][m] = ek;
                ++m;
            }
        // MLS_SF_d with compute_AB_d
        double w[FRIBO_DIMSD][3], pk[3][FRIBO_DIMSD], A[3][3][3];
        for (int i = 0; i < FRIBO_DIMSD; ++i) {
            pk[0][i] = 1.0;
            pk[1][i] = sdc[0][i];
            pk[2][i] = sdc[1][i];
            double w1, d1, w2, d2;
            weight1d_d(probe[0] - sdc[0][i], sdd[0][i], F.wfun, F.radiusSD, w1, d1);
            weight1d_d(probe[1] - sdc[1][i], sdd[1][i], F.wfun, F.radiusSD, w2, d2);
            w[i][0] = w1 * w2;
            w[i][1] = w2 * d1;
            w[i][2] = w1 * d2;
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int l = 0; l < 3; ++l) A[a][b][l] = 0.0;
        for (int i = 0; i < FRIBO_DIMSD; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double pp = pk[a][i] * pk[b][i];
                    for (int l = 0; l < 3; ++l) A[a][b][l] = A[a][b][l] + w[i][l] * pp;
                }
        double Aa[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Aa[a][b] = A[a][b][0];
        double gam[3][3];
        const double g0[3] = {1.0, probe[0], probe[1]};
        double x[3] = {0.0, 0.0, 0.0};
        gesolve3(Aa, g0, x);
        for (int a = 0; a < 3; ++a) gam[a][0] = x[a];
        for (int l = 1; l <= 2; ++l) {
            double C[3];
            for (int a = 0; a < 3; ++a) {
                double mv = 0.0;
                for (int b = 0; b < 3; ++b) mv = mv + A[a][b][l] * gam[b][0];
                C[a] = ((a == l) ? 1.0 : 0.0) - mv;
            }
            x[0] = x[1] = x[2] = 0.0;
            gesolve3(Aa, C, x);
            for (int a = 0; a < 3; ++a) gam[a][l] = x[a];
        }
        double ph[3][FRIBO_DIMSD], sumphi = 0.0;
        for (int i = 0; i < FRIBO_DIMSD; ++i) {
            ph[0][i] = 0.0;
            ph[1][i] = 0.0;
            ph[2][i] = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double b0 = pk[a][i] * w[i][0], b1 = pk[a][i] * w[i][1], b2 = pk[a][i] * w[i][2];
                ph[0][i] = ph[0][i] + gam[a][0] * b0;
                ph[1][i] = ph[1][i] + gam[a][1] * b0 + gam[a][0] * b1;
                ph[2][i] = ph[2][i] + gam[a][2] * b0 + gam[a][0] * b2;
            }
        }
        for (int i = 0; i < FRIBO_DIMSD; ++i) sumphi = sumphi + ph[0][i];
        if (fabs(1.0 - fabs(sumphi)) > 1.0e-10) atomicCAS(&S->err, 0, FRIBO_E_PARTITION_UNITY);
        // getpbup
        double* uu = u[idir - 1];
        for (int i = 0; i < 7; ++i) uu[i] = 0.0;
        const double* f = (idir == 1) ? F.q2 : F.q3;
        for (int i = 0; i < FRIBO_DIMSD; ++i) {
            const int cy = eul[0][i], cz = eul[1][i];
            double fc, fjm, fkm;
            if (idir == 3) {
                fc = prat(F, S, cy, cz);
                fjm = prat(F, S, cy - 1, cz);
                fkm = prat(F, S, cy, cz - 1);
            } else {
                fc = fat(F, f, cy, cz);
                fjm = fat(F, f, cy - 1, cz);
                fkm = fat(F, f, cy, cz - 1);
            }
            const double uk0 = fc;
            const double uk1 = (fc - fjm) * F.dx2 / F.g2rc[cy];
            const double uk2 = (fc - fkm) * F.dx3 / F.g3rc[cz];
            for (int q = 0; q < 3; ++q) uu[q] = uu[q] + ph[q][i] * uk0;
            uu[3] = uu[3] + ph[1][i] * uk1;
            uu[4] = uu[4] + ph[1][i] * uk2;
            uu[5] = uu[5] + ph[2][i] * uk1;
            uu[6] = uu[6] + ph[2][i] * uk2;
        }
    }
    const double* ucy = u[0];
    const double* ucz = u[1];
    const double* pc = u[2];
    const double veln_md = ucy[0] * nor[0] + ucz[0] * nor[1];
    const double vt0 = ucy[0] - veln_md * nor[0], vt1 = ucz[0] - veln_md * nor[1];
    const double velt_md = sqrt(vt0 * vt0 + vt1 * vt1);
    LW(n, 0) = 1.0;
    if (velt_md != 0.0) {
        const double nrm = sqrt(vt0 * vt0 + vt1 * vt1);
        LW(n, 1) = 1.0;
        LW(n, 2) = vt0 / nrm;
        LW(n, 3) = vt1 / nrm;
    }
    double dpdnw = LMACC(n, 1) * nor[0] + LMACC(n, 2) * nor[1];
    dpdnw = -1.0 * dpdnw;
    const double dpdne = pc[1] * nor[0] + pc[2] * nor[1];
    F.press[n - 1] = pc[0] - 0.5 * (dpdnw + dpdne) * hpr;
    for (int q = 1; q < 7; ++q) {
        LW(n, 3 + q) = ucy[q];
        LW(n, 9 + q) = ucz[q];
    }
    const double nory = nor[0], norz = nor[1];
    LW(n, 16) = (pc[3] * nory * nory + pc[4] * nory * norz + pc[5] * norz * nory + pc[6] * norz * norz) * 1.0 / F.visc;
}

// Tangent of every marker: its own, else the last one set before it
// (loads_transfer.F90:94-96), by one wave with a "last set" scan over lane
// chunks; then the friction stress (loads_transfer.F90:111-128).
__global__ __launch_bounds__(64) void k_loads_tau(const FluidDev F, double hpr)
{
    DevScalars* S = F.S;
    if (S->err) return;
    const int lane = threadIdx.x, nlm = F.nlm;
    const int per = (nlm + 63) / 64;
    const int n0 = lane * per + 1, n1 = min(nlm, n0 + per - 1);
    bool has = false;
    double ty = 0.0, tz = 0.0;
    for (int n = n0; n <= n1; ++n)
        if (LW(n, 0) != 0.0 && LW(n, 1) != 0.0) {
            has = true;
            ty = LW(n, 2);
            tz = LW(n, 3);
        }
    // inclusive scan of (has, t) with "later set wins", then shift by one lane
    for (int h = 1; h < 64; h <<= 1) {
        const int oh = __shfl_up((int)has, h, 64);
        const double oy = __shfl_up(ty, h, 64), oz = __shfl_up(tz, h, 64);
        if (lane >= h && !has && oh) {
            has = true;
            ty = oy;
            tz = oz;
        }
    }
    int ih = __shfl_up((int)has, 1, 64);
    double cy = __shfl_up(ty, 1, 64), cz = __shfl_up(tz, 1, 64);
    if (lane == 0 || !ih) {
        cy = 0.0;
        cz = 0.0;
    }
    const double visc = F.visc;
    for (int n = n0; n <= n1; ++n) {
        F.tau[(size_t)(n - 1) + nlm] = 0.0;
        if (LW(n, 0) == 0.0) {
            F.tau[n - 1] = 0.0;
            continue;
        }
        if (LW(n, 1) != 0.0) {
            cy = LW(n, 2);
            cz = LW(n, 3);
        }
        const double nory = LMNOR(n, 1), norz = LMNOR(n, 2), tany = cy, tanz = cz;
        const double u1 = LW(n, 4), u2 = LW(n, 5), u3 = LW(n, 6), u4 = LW(n, 7), u5 = LW(n, 8), u6 = LW(n, 9);
        const double v1 = LW(n, 10), v2 = LW(n, 11), v3 = LW(n, 12), v4 = LW(n, 13), v5 = LW(n, 14), v6 = LW(n, 15);
        const double d3ucdn3 = LW(n, 16);
        const double ducdn = (u1 * tany * nory + u2 * tany * norz + v1 * tanz * nory + v2 * tanz * norz);
        const double d2ucdn2 = (u3 * tany * nory * nory + u4 * tany * nory * norz + u5 * tany * norz * nory +
                                u6 * tany * norz * norz + v3 * tanz * nory * nory + v4 * tanz * nory * norz +
                                v5 * tanz * norz * nory + v6 * tanz * norz * norz);
        const double ducdnm = ducdn - d2ucdn2 * hpr - 0.5 * d3ucdn3 * (hpr * hpr);
        F.tau[n - 1] = visc * ducdnm * tany;
        F.tau[(size_t)(n - 1) + nlm] = visc * ducdnm * tanz;
    }
}

void launch_loads(const FluidDev& F, double hpr, hipStream_t st)
{
    if (F.nlm <= 0) return;
    hipLaunchKernelGGL(k_loads, dim3((F.nlm + 63) / 64), dim3(64), 0, st, F, hpr);
    hipLaunchKernelGGL(k_loads_tau, dim3(1), dim3(64), 0, st, F, hpr);
}

// ckbcs.F90:5-86: velocity error at the markers after the step and the norms
// sqrt(sum err^2)/nlm (fixed-order tree sum); nb[0..1] receive the norms.
__global__ __launch_bounds__(1024) void k_checkbcs(const FluidDev F, double* nb)
{
    __shared__ double sv[1024], sw[1024];
    const int t = threadIdx.x;
    double av = 0.0, aw = 0.0;
    for (int np = 1 + t; np <= F.nlm; np += 1024) {
        double ev = 0.0, ew = 0.0;
        const double z = LMCC(np, 2);
        if (F.zz[1] < z && z < F.zz[F.n3]) {
            double ucy = 0.0, ucz = 0.0;
            for (int m = 1; m <= FRIBO_DIMSD; ++m)
                ucy = ucy + PHIPRED(np, 1, m) * F.q2[fidx(F, EULINS(np, 1, 1, m), EULINS(np, 1, 2, m))];
            ev = LMVEL(np, 1) - ucy;
            for (int m = 1; m <= FRIBO_DIMSD; ++m)
                ucz = ucz + PHIPRED(np, 2, m) * F.q3[fidx(F, EULINS(np, 2, 1, m), EULINS(np, 2, 2, m))];
            ew = LMVEL(np, 2) - ucz;
        }
        F.bcerr[np - 1] = ev;
        F.bcerr[(size_t)(np - 1) + F.nlm] = ew;
        av = av + ev * ev;
        aw = aw + ew * ew;
    }
    sv[t] = av;
    sw[t] = aw;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) {
            sv[t] = sv[t] + sv[t + h];
            sw[t] = sw[t] + sw[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        nb[0] = sqrt(sv[0]) / (double)F.nlm;
        nb[1] = sqrt(sw[0]) / (double)F.nlm;
    }
}

void launch_checkbcs(const FluidDev& F, double* nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_checkbcs, dim3(1), dim3(1024), 0, st, F, nb);
}
