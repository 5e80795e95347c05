// Batched launches of the step's kernels over many independent instances
// (uneven grids) with one launch per stage.
//
// Every kernel body takes its block coordinates from a VBlock instead of
// blockIdx / gridDim.  The single-instance kernel fills it from the hardware
// registers; the batched kernel gets a 1-D grid that concatenates the grids
// of all instances (instance i owns blocks [off[i], off[i+1]) laid out as its
// own dims[i], x fastest), finds its instance by a binary search over off
// and runs the same body on that instance's FluidDev.  An instance's blocks
// therefore do exactly the work of its single-instance launch.
#pragma once

#include "fribo_internal.h"

#include <vector>

struct VBlock {
    int x, y, z;
    int gx, gy, gz;
};

__device__ __forceinline__ VBlock vb_hw()
{
    return VBlock{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// Launch table of a batched launch; T is the per-instance payload (the
// instance's FluidDev, or a GEMM problem).
template <class T>
struct BGrid {
    const T* item;       // [n]
    const int* off;      // [n+1] first block of each instance
    const int* dims;     // [n][3] the instance's own grid
    int n;
};

// Instance of the current block (the last one whose first block is <= it:
// instances with no blocks are skipped) and the block's coordinates in it.
template <class T>
__device__ __forceinline__ const T* bg_block(const BGrid<T>& B, VBlock& v)
{
    const int b = (int)blockIdx.x;
    int lo = 0, hi = B.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const int l = b - B.off[lo];
    v.gx = B.dims[3 * lo];
    v.gy = B.dims[3 * lo + 1];
    v.gz = B.dims[3 * lo + 2];
    v.x = l % v.gx;
    v.y = (l / v.gx) % v.gy;
    v.z = l / (v.gx * v.gy);
    return B.item + lo;
}

// bg_block for a launch of bg_total rounded up to a multiple of 8 blocks:
// the hardware hands blocks to the 8 XCDs round-robin, so block b takes
// virtual block (b % 8) * (blocks / 8) + b / 8 and each XCD works through a
// contiguous range of the instances' tiles (neighbouring tiles, which share
// the cache lines at their edges, meet in one L2).  nullptr for the padding
// blocks.
template <class T>
__device__ __forceinline__ const T* bg_block_xcd(const BGrid<T>& B, VBlock& v)
{
    const int total = B.off[B.n];
    const int nb = (int)gridDim.x;
    const int b = ((int)blockIdx.x & 7) * (nb >> 3) + ((int)blockIdx.x >> 3);
    if (b >= total) return nullptr;
    int lo = 0, hi = B.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const int l = b - B.off[lo];
    v.gx = B.dims[3 * lo];
    v.gy = B.dims[3 * lo + 1];
    v.gz = B.dims[3 * lo + 2];
    v.x = l % v.gx;
    v.y = (l / v.gx) % v.gy;
    v.z = l / (v.gx * v.gy);
    return B.item + lo;
}

// GEMM problem of one instance (the arguments of a single-instance launch).
struct GemmProb {
    int M, N, K, lda, ldb, ldc;
    const double* A;
    const double* B;
    double* C;
    const DevScalars* S;
    const double* zg;
    double* gpart;
    YFuse yf;
    int ks;   // mirror form: split point of the block-diagonal A (0: plain)
};

// Host side of a batch, handed to the batched launchers.  hF are the
// instances' FluidDev (host copies), dF the same array on the device.
// grid() returns the launch table of a stage (cached under key, rebuilt
// when the instances' grids change); gemm() likewise for GEMM problems.
struct BatchSet {
    int n = 0;
    std::vector<FluidDev> hF;
    FluidDev* dF = nullptr;
    BGrid<FluidDev> (*grid_fn)(BatchSet*, int key, const std::vector<int>& dims) = nullptr;
    BGrid<GemmProb> (*gemm_fn)(BatchSet*, int key, const std::vector<GemmProb>& probs,
                               const std::vector<int>& dims) = nullptr;
    void* owner = nullptr;
    // choices shared by all instances (largest size class over the batch)
    int gemm_cfg = 0;        // 0: 32x32 tiles, 1: 64x64, 2: 96x128, 3: direct-to-LDS 96x128, 4: direct-to-LDS 96x64
    bool ysolve_fused = false;

    BGrid<FluidDev> grid(int key, const std::vector<int>& dims) { return grid_fn(this, key, dims); }
    BGrid<GemmProb> gemm(int key, const std::vector<GemmProb>& p, const std::vector<int>& dims)
    {
        return gemm_fn(this, key, p, dims);
    }
};

// Per-instance grids of a stage: fn(F) -> dim3 for every instance.
template <class Fn>
inline std::vector<int> bdims(const BatchSet& B, Fn fn)
{
    std::vector<int> d(3 * (size_t)B.n);
    for (int i = 0; i < B.n; ++i) {
        const dim3 g = fn(B.hF[i]);
        d[3 * i] = (int)g.x;
        d[3 * i + 1] = (int)g.y;
        d[3 * i + 2] = (int)g.z;
    }
    return d;
}

// Total blocks of a table built from dims.
inline int bg_total(const std::vector<int>& dims)
{
    long t = 0;
    for (size_t i = 0; i + 2 < dims.size(); i += 3) t += (long)dims[i] * dims[i + 1] * dims[i + 2];
    return (int)t;
}

// One batched launch over the table of `key` built from dims; nothing is
// launched when no instance has blocks.
template <class Kern, class... Args>
inline void blaunch(BatchSet& B, int key, const std::vector<int>& dims, Kern kern, dim3 block, size_t lds,
                    hipStream_t st, Args... args)
{
    const int tot = bg_total(dims);
    if (tot <= 0) return;
    hipLaunchKernelGGL(kern, dim3(tot), block, lds, st, B.grid(key, dims), args...);
}

// blaunch for kernels that look their block up with bg_block_xcd.
template <class Kern, class... Args>
inline void blaunch_xcd(BatchSet& B, int key, const std::vector<int>& dims, Kern kern, dim3 block, size_t lds,
                        hipStream_t st, Args... args)
{
    const int tot = bg_total(dims);
    if (tot <= 0) return;
    hipLaunchKernelGGL(kern, dim3((tot + 7) / 8 * 8), block, lds, st, B.grid(key, dims), args...);
}

// Stage keys of the launch tables.
enum {
    BK_CFL, BK_BC, BK_RHS, BK_DIV, BK_UPD, BK_MEAN, BK_MLS_IDX, BK_MLS_SUP, BK_MLS_ORDER, BK_MLS_INTERP,
    BK_MLS_SPREAD, BK_JH, BK_JP, BK_K1, BK_K2, BK_K3, BK_KF, BK_Y1, BK_Y2, BK_Y3, BK_GEMM_F, BK_GEMM_B, BK_IO,
    BK_COUNT
};

void launch_cfl_b(BatchSet& B, hipStream_t st);
void launch_bc_b(BatchSet& B, int n, bool mean, hipStream_t st);
void launch_rhs_b(BatchSet& B, hipStream_t st);
void launch_div_b(BatchSet& B, hipStream_t st);
void launch_upd_b(BatchSet& B, hipStream_t st);
void launch_mean_b(BatchSet& B, hipStream_t st);
void launch_mls_setup_b(BatchSet& B, hipStream_t st);
void launch_mls_force_b(BatchSet& B, hipStream_t st);
void launch_jsolve_helm_b(BatchSet& B, hipStream_t st);
void launch_jsolve_poisson_b(BatchSet& B, hipStream_t st);
void launch_ksolve_b(BatchSet& B, hipStream_t st);
void launch_ysolve_b(BatchSet& B, hipStream_t st, bool first_done);
// the first (chunk) pass of the batched per-mode z-lines alone
void launch_ysolve1_b(BatchSet& B, hipStream_t st);
// transforms of the Poisson solve: forward (fwd) or backward, the axis of
// the instances (all alike); returns true when the fused chunk pass ran
bool launch_poisson_gemm_b(BatchSet& B, bool fwd, hipStream_t st);
int jsolve_class_h(const FluidDev& F);
int gemm_cfg_mask(const FluidDev& F);
int gemm_cfg_parts(const FluidDev& F, int cfg);
bool gemm_fuse_ok(const FluidDev& F, int cfg);
int jsolve_stride_for(const FluidDev& F, int mh);
// true when every instance can use the counting-sort marker order
bool mls_batchable(const FluidDev& F);
