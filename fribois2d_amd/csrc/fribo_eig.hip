// Host-side eigen-decomposition of the y pressure operator (set-up only).
//
// The Poisson solve of phcalc_outflow.F90:4-71 diagonalises the 1-D Neumann
// operator along z (zmm, zmmt, wr from phini.F90:35-99) and solves one
// tridiagonal system along y per z-mode.  The same discrete problem can be
// diagonalised along y instead (one tridiagonal system along z per y-mode);
// the dense transform then costs 2*n3m*n2m^2 instead of 2*n2m*n3m^2 flops,
// which is cheaper whenever n2m < n3m.  This file builds the y eigen-system.
//
// A_y (rows amphj, acphj, apphj; Neumann at both ends) is similar to a
// symmetric tridiagonal matrix: with D = diag(d), d_1 = 1,
// (d_{j+1}/d_j)^2 = A(j+1,j)/A(j,j+1), S = D^-1 A_y D has off-diagonal
// sqrt(A(j,j+1) A(j+1,j)).  S = Q diag(w) Q^T by the implicit QL method with
// Wilkinson shifts (Bowdler, Martin, Reinsch and Wilkinson, "The QR and QL
// algorithms for symmetric matrices", Numer. Math. 11 (1968); EISPACK tql2),
// so A_y = (D Q) diag(w) (Q^T D^-1): Y = D Q, Yinv = Q^T D^-1.
#include "fribo_gpu.h"
#include "fribo_internal.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

// Symmetric tridiagonal eigen-problem by implicit QL.  d[0..n-1] diagonal
// (overwritten by the eigenvalues), e[0..n-2] off-diagonal (destroyed),
// z (n x n, column-major, initialised to I) receives the orthonormal
// eigenvectors as columns.  Returns 0, or the index of an eigenvalue that
// did not converge in 60 iterations.  The iteration on (d, e) does not
// depend on z, so the plane rotations are recorded first and then replayed
// on z by several threads, each owning a block of rows (rotations act on
// column pairs, so rows are independent).
struct Rot {
    int i;
    double s, c;
};

int tridiag_ql(int n, double* d, double* e, double* z)
{
    std::vector<double> ee(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) ee[i] = e[i];
    std::vector<Rot> rots;
    rots.reserve((size_t)n * n);
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(ee[m]) + dd == dd) break;   // negligible in FP64
            }
            if (m == l) break;
            if (++iter > 60) return l + 1;
            // Wilkinson shift from the leading 2x2 block
            double g = (d[l + 1] - d[l]) / (2.0 * ee[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + ee[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool deflated = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * ee[i];
                const double b = c * ee[i];
                r = std::hypot(f, g);
                ee[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    ee[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                rots.push_back({i, s, c});
            }
            if (deflated) continue;
            d[l] -= p;
            ee[l] = g;
            ee[m] = 0.0;
        }
    }
    // replay the rotations on z (columns i, i+1), rows split over threads
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt < 1 ? 1 : (nt > 8 ? 8 : nt);
    if ((size_t)n < 64 * (size_t)nt) nt = 1;
    auto work = [&](int k0, int k1) {
        for (const Rot& q : rots) {
            double* zi = z + (size_t)q.i * n;
            double* zj = zi + n;
            const double s = q.s, c = q.c;
            for (int k = k0; k < k1; ++k) {
                const double f = zj[k];
                zj[k] = s * zi[k] + c * f;
                zi[k] = c * zi[k] - s * f;
            }
        }
    };
    std::vector<std::thread> th;
    const int blk = (n + (int)nt - 1) / (int)nt;
    for (unsigned t = 1; t < nt; ++t) {
        const int k0 = (int)t * blk, k1 = std::min(n, k0 + blk);
        if (k0 < k1) th.emplace_back(work, k0, k1);
    }
    work(0, std::min(n, blk));
    for (auto& x : th) x.join();
    return 0;
}

// Y = D Q and Yinv = Q^T D^-1 from the orthonormal eigenvectors Q (n x n,
// column-major) of the symmetrised operator: thread j of column block m.
__global__ __launch_bounds__(256) void k_eig_assemble(int n, int ld, const double* __restrict__ Q,
                                                      const double* __restrict__ dg, double* __restrict__ Y,
                                                      double* __restrict__ Yinv)
{
    const int j = blockIdx.x * 256 + threadIdx.x, m = blockIdx.y;
    if (j >= n) return;
    const double q = Q[(size_t)m * n + j];
    Y[(size_t)m * ld + j] = dg[j] * q;
    Yinv[(size_t)j * ld + m] = q / dg[j];
}

// The decomposition on the device: the symmetrised tridiagonal problem by
// rocSOLVER's divide-and-conquer eigensolver (STEDC: eigenvalues ascending,
// orthonormal eigenvectors), the scaling back to Y and Yinv in a kernel.
// Host arrays in and out, like the host path.
static int eig_device(int n, const std::vector<double>& d, const std::vector<double>& e,
                      const std::vector<double>& dg, int ld, double* Y, double* Yinv, double* wy)
{
    double *dD = nullptr, *dE = nullptr, *dQ = nullptr, *dG = nullptr, *dY = nullptr, *dYi = nullptr;
    rocblas_int* dinfo = nullptr;
    rocblas_handle hb = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    rocblas_int info = 0;
    const size_t nn = (size_t)n * n, ln = (size_t)ld * n;
    auto ok = [&](hipError_t he) {
        if (he != hipSuccess && !rc) rc = FRIBO_ERR_HIP;
        return rc == 0;
    };
    if (ok(hipStreamCreate(&st)) && ok(hipMalloc(&dD, n * sizeof(double))) &&
        ok(hipMalloc(&dE, std::max(n - 1, 1) * sizeof(double))) && ok(hipMalloc(&dQ, nn * sizeof(double))) &&
        ok(hipMalloc(&dG, n * sizeof(double))) && ok(hipMalloc(&dY, ln * sizeof(double))) &&
        ok(hipMalloc(&dYi, ln * sizeof(double))) && ok(hipMalloc(&dinfo, sizeof(rocblas_int))) &&
        ok(hipMemcpyAsync(dD, d.data(), n * sizeof(double), hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(dE, e.data(), std::max(n - 1, 1) * sizeof(double), hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(dG, dg.data(), n * sizeof(double), hipMemcpyHostToDevice, st)) &&
        ok(hipMemsetAsync(dY, 0, ln * sizeof(double), st)) && ok(hipMemsetAsync(dYi, 0, ln * sizeof(double), st))) {
        if (rocblas_create_handle(&hb) != rocblas_status_success || rocblas_set_stream(hb, st) != rocblas_status_success ||
            rocsolver_dstedc(hb, rocblas_evect_tridiagonal, n, dD, dE, dQ, n, dinfo) != rocblas_status_success)
            rc = FRIBO_ERR_HIP;
        if (!rc) {
            hipLaunchKernelGGL(k_eig_assemble, dim3((n + 255) / 256, n), dim3(256), 0, st, n, ld, dQ, dG, dY, dYi);
            if (ok(hipGetLastError()) && ok(hipMemcpyAsync(&info, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st)) &&
                ok(hipMemcpyAsync(wy, dD, n * sizeof(double), hipMemcpyDeviceToHost, st)) &&
                ok(hipMemcpyAsync(Y, dY, ln * sizeof(double), hipMemcpyDeviceToHost, st)) &&
                ok(hipMemcpyAsync(Yinv, dYi, ln * sizeof(double), hipMemcpyDeviceToHost, st)) &&
                ok(hipStreamSynchronize(st)) && info != 0)
                rc = FRIBO_ERR_HIP;   // STEDC did not converge
        }
    }
    if (hb) rocblas_destroy_handle(hb);
    if (st) hipStreamSynchronize(st);
    hipFree(dD);
    hipFree(dE);
    hipFree(dQ);
    hipFree(dG);
    hipFree(dY);
    hipFree(dYi);
    hipFree(dinfo);
    if (st) hipStreamDestroy(st);
    return rc;
}

// Y (eigenvectors of A_y as columns), Yinv = Y^-1 and the eigenvalues wy of
// the n x n operator with rows (am, ac, ap) (0-based, am[0] = ap[n-1] = 0).
// Y and Yinv are column-major with leading dimension ld.  device: solve the
// symmetrised problem on the GPU (eig_device) instead of the host QL.
int ytransform_build(int n, const double* am, const double* ac, const double* ap, int ld, double* Y, double* Yinv,
                     double* wy, bool device)
{
    std::vector<double> d(n), e(n > 1 ? n - 1 : 1), dg(n), q((size_t)n * n, 0.0);
    dg[0] = 1.0;
    for (int j = 0; j + 1 < n; ++j) {
        const double up = ap[j], lo = am[j + 1];
        if (!(up > 0.0) || !(lo > 0.0)) return -1;   // not similarity-symmetrisable
        e[j] = std::sqrt(up * lo);
        dg[j + 1] = dg[j] * std::sqrt(lo / up);
    }
    for (int j = 0; j < n; ++j) d[j] = ac[j];
    if (device) return eig_device(n, d, e, dg, ld, Y, Yinv, wy);
    for (int j = 0; j < n; ++j) q[(size_t)j * n + j] = 1.0;
    const int rc = tridiag_ql(n, d.data(), e.data(), q.data());
    if (rc) return rc;
    for (int m = 0; m < n; ++m) {
        wy[m] = d[m];
        for (int j = 0; j < n; ++j) {
            const double qv = q[(size_t)m * n + j];   // Q(j, m)
            Y[(size_t)m * ld + j] = dg[j] * qv;       // Y(j, m) = d_j Q(j, m)
            Yinv[(size_t)j * ld + m] = qv / dg[j];    // Yinv(m, j) = Q(j, m) / d_j
        }
    }
    return 0;
}

// Mirror form of the y eigen-system.  When the grid is mirror-symmetric
// about the channel centre (am[j] = ap[n-1-j], ac[j] = ac[n-1-j] to within
// tol * max|ac|) and n = 2h is even, the operator commutes with the
// reflection J (j -> n-1-j), so every eigenvector is even ([u; Ju]) or odd
// ([w; -Ju]).  Even modes are the eigenvectors u of the top half with the
// centre coupling added to its last diagonal entry, odd modes those with it
// subtracted: two h x h problems built from the top half's coefficients
// (the bottom half's are their mirror image).  With the pair butterfly
// (x_j, x_{n-1-j}) -> (x_j + x_{n-1-j}, x_j - x_{n-1-j}), j < h, applied to
// the field before the forward transform and to the result of the backward
// one, both transforms become block-diagonal with two h x h blocks, half
// the flops of the dense form:
//   forward  (modes i, positions kk): Ysi(i, kk) = u_i(kk) / (sqrt2 d_kk)
//            for i, kk < h; Ysi(h + o, n-1-j) = w_o(j) / (sqrt2 d_j);
//   backward (positions i, modes kk): Ys(i, kk) = d_i u_kk(i) / sqrt2 for
//            i, kk < h; Ys(n-1-j, h + o) = d_j w_o(j) / sqrt2;
// every other entry 0.  wys holds the even block's eigenvalues, then the
// odd block's.  Returns -1 when the operator is not mirror-symmetric.
int ytransform_mirror(int n, const double* am, const double* ac, const double* ap, int ld, double* Ys, double* Ysi,
                      double* wys, double tol)
{
    if (n < 4 || (n & 1)) return -1;
    const int h = n / 2;
    double amax = 0.0;
    for (int j = 0; j < n; ++j) amax = std::max(amax, std::fabs(ac[j]));
    if (!(amax > 0.0)) return -1;
    for (int j = 0; j < n; ++j)
        if (!(std::fabs(am[j] - ap[n - 1 - j]) <= tol * amax) || !(std::fabs(ac[j] - ac[n - 1 - j]) <= tol * amax))
            return -1;
    std::vector<double> dg(h), e(h), de(h), doo(h), ee(h), eo(h);
    dg[0] = 1.0;
    for (int j = 0; j < h; ++j) {
        const double up = ap[j], lo = am[j + 1];   // j = h-1: the coupling across the centre
        if (!(up > 0.0) || !(lo > 0.0)) return -1;
        e[j] = std::sqrt(up * lo);
        if (j + 1 < h) dg[j + 1] = dg[j] * std::sqrt(lo / up);
    }
    for (int j = 0; j < h; ++j) {
        de[j] = ac[j];
        doo[j] = ac[j];
        ee[j] = e[j];
        eo[j] = e[j];
    }
    de[h - 1] = de[h - 1] + e[h - 1];
    doo[h - 1] = doo[h - 1] - e[h - 1];
    std::vector<double> qe((size_t)h * h, 0.0), qo((size_t)h * h, 0.0);
    for (int j = 0; j < h; ++j) qe[(size_t)j * h + j] = qo[(size_t)j * h + j] = 1.0;
    if (tridiag_ql(h, de.data(), ee.data(), qe.data()) || tridiag_ql(h, doo.data(), eo.data(), qo.data())) return -1;
    const double r2 = std::sqrt(0.5);
    std::fill(Ys, Ys + (size_t)ld * n, 0.0);
    std::fill(Ysi, Ysi + (size_t)ld * n, 0.0);
    for (int m = 0; m < h; ++m) {
        wys[m] = de[m];
        wys[h + m] = doo[m];
        for (int j = 0; j < h; ++j) {
            const double u = qe[(size_t)m * h + j], w = qo[(size_t)m * h + j];   // u_m(j), w_m(j)
            Ysi[(size_t)j * ld + m] = u * r2 / dg[j];
            Ysi[(size_t)(n - 1 - j) * ld + h + m] = w * r2 / dg[j];
            Ys[(size_t)m * ld + j] = dg[j] * u * r2;
            Ys[(size_t)(h + m) * ld + n - 1 - j] = dg[j] * w * r2;
        }
    }
    return 0;
}

extern "C" int fribo_test_ymirror(int32_t n, const double* am, const double* ac, const double* ap, double tol,
                                  double* Ys, double* Ysi, double* wys)
{
    if (n <= 0 || !am || !ac || !ap || !Ys || !Ysi || !wys) return 5;
    return ytransform_mirror(n, am, ac, ap, n, Ys, Ysi, wys, tol) == 0 ? 0 : 7;
}

extern "C" int fribo_test_ytransform(int32_t n, const double* am, const double* ac, const double* ap, double* Y,
                                     double* Yinv, double* wy)
{
    if (n <= 0 || !am || !ac || !ap || !Y || !Yinv || !wy) return 5;
    return ytransform_build(n, am, ac, ap, n, Y, Yinv, wy) == 0 ? 0 : 7;
}

// phini (phini.F90:4-99) without DGEEV/DGESV: the z pressure operator of
// tridiag_matrices (phini.F90:121-142, Neumann at both ends) from the node
// coordinates zz(1:n3), decomposed like the y operator above.  zmm
// (eigenvectors as columns) and zmmt = zmm^-1 are n3m x n3m with leading
// dimension ld; wr(n3m) the eigenvalues.  The eigen-pairs are those of
// DGEEV up to order and scaling, which the Poisson solve does not see.
int phini_sym(int n3, const double* zz, double alx3, int ld, double* zmm, double* zmmt, double* wr, bool device)
{
    const int n3m = n3 - 1;
    if (n3m < 2 || ld < n3m) return -1;
    const double dx3 = 1.0 / (alx3 / (double)n3m), dx3q = dx3 * dx3;
    std::vector<double> g3rm(n3m + 2), g3rc(n3 + 2), am(n3m), ac(n3m), ap(n3m);
    // cordin.F90:197-215 (1-based)
    for (int k = 1; k <= n3m; ++k) g3rm[k] = (zz[k] - zz[k - 1]) * dx3;
    for (int k = 2; k <= n3m; ++k) g3rc[k] = (zz[k] - zz[k - 2]) * dx3 * 0.5;
    g3rc[1] = (zz[1] - zz[0]) * dx3;
    g3rc[n3] = (zz[n3 - 1] - zz[n3m - 1]) * dx3;
    for (int kc = 1; kc <= n3m; ++kc) {
        const int kmc = (kc == 1) ? 0 : 1, kpc = (kc == n3m) ? 0 : 1;
        const int kp = (kc == n3m) ? kc : kc + 1;
        const double a33icc = kmc * dx3q / g3rc[kc], a33icp = kpc * dx3q / g3rc[kp], ug = 1.0 / g3rm[kc];
        am[kc - 1] = a33icc * ug;
        ap[kc - 1] = a33icp * ug;
        ac[kc - 1] = -(am[kc - 1] + ap[kc - 1]);
    }
    return ytransform_build(n3m, am.data(), ac.data(), ap.data(), ld, zmm, zmmt, wr, device);
}

extern "C" int fribo_phini(int32_t n3, const double* zz, double alx3, int32_t ld, double* zmm, double* zmmt,
                           double* wr)
{
    if (n3 < 3 || !zz || !zmm || !zmmt || !wr || ld < n3 - 1 || !(alx3 > 0.0)) return 5;
    return phini_sym(n3, zz, alx3, ld, zmm, zmmt, wr) == 0 ? 0 : 7;
}

// phini on the GPU: the same operator and similarity, eigen-pairs by
// rocSOLVER's STEDC (eigenvalues in increasing order).
extern "C" int fribo_phini_device(int32_t n3, const double* zz, double alx3, int32_t ld, double* zmm, double* zmmt,
                                  double* wr)
{
    if (n3 < 3 || !zz || !zmm || !zmmt || !wr || ld < n3 - 1 || !(alx3 > 0.0)) return FRIBO_ERR_ARG;
    const int rc = phini_sym(n3, zz, alx3, ld, zmm, zmmt, wr, true);
    return rc == 0 ? 0 : (rc == FRIBO_ERR_HIP ? FRIBO_ERR_HIP : FRIBO_ERR_LIMIT);
}
