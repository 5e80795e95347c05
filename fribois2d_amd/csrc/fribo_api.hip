// C ABI of libfribo_gpu (include/fribo_gpu.h): set-up, state transfer and the
// per-step orchestration that replaces fluid_tsch.F90:25-166 on one GPU.
//
// One HIP stream per handle carries every kernel of a step; the only host
// synchronisation per step is the final read-back of dt / cflm / the error
// flag (the values the Fortran driver needs before rigid() and ttime()).
#include "fribo_internal.h"
#include "fribo_batch.h"
#include "../../include/fribo_gpu.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

static thread_local std::string g_err = "";

static int set_err(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return set_err(FRIBO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static const char* kProfNames[FRIBO_NPROF] = {"cfl", "bc", "rhs", "mls_setup", "mls_force", "jsolve_helm",
                                              "ksolve", "div", "gemm", "jsolve_poisson", "upd", "mean"};
enum { PC_CFL, PC_BC, PC_RHS, PC_MLS_SETUP, PC_MLS_FORCE, PC_JH, PC_K, PC_DIV, PC_GEMM, PC_JP, PC_UPD, PC_MEAN };

struct fribo_ctx {
    fribo_params p;
    FluidDev F;
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    // host copies of the 1-D metric arrays (1-based, size n+2)
    std::vector<double> h1[32];
    std::vector<double> lmcc;   // host copy for the domain check
    int nlm_alloc = 0;
    std::vector<void*> mk_allocs;
    // marker upload: lmcc, lmvel, lmA staged together in pinned memory and
    // sent in one copy; ev_mk marks the copy so the staging is reused only
    // after it has been read
    double* mk_stage = nullptr;
    hipEvent_t ev_mk = nullptr;
    DevScalars hs;
    // profiling
    bool profiling = false;
    unsigned prof_mask = ~0u;   // kernel classes bracketed by events
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    struct Rec { int cls; hipEvent_t a, b; };
    std::vector<Rec> pending;
    double prof_ms[FRIBO_NPROF] = {0};
    int prof_n[FRIBO_NPROF] = {0};
    size_t lds_limit = 65536;
    // step inputs staged in pinned memory: {dt, dt_o, ns, err=0}
    struct StepIn { double dt, dto; int ns, err; };
    StepIn* in = nullptr;
    DevScalars* hsp = nullptr;   // pinned landing buffer of the scalars read back after a step
    hipEvent_t ev_rb = nullptr;  // the read-back copy of those scalars
    // k_cfl of the next step already enqueued behind the read-back of this
    // one (it reads only q2/q3, so it runs while the host prepares the next
    // step); cleared by anything that changes the velocities in between
    bool cfl_ready = false;
    // captured step (hipGraph): rebuilt when the marker buffers change
    bool use_graph = false;
    bool yavail = false;     // transform along y set up
    double dmin = 0.0;       // smallest node spacing of the grid (cordin.F90:217-231)
    double* nb = nullptr;    // device: the two checkbcs norms
    double* dpart = nullptr; // device: diagnostics partials [8][256] and results [9]
    // rollback copy of the prognostic state for iterated coupling
    double* snap = nullptr;
    bool snap_valid = false;
    int vort_k = 0;          // plane of getintvort's z = 18.5 line (0: none)
    hipGraphExec_t gexec = nullptr;
    int graph_nlm = -1;
    int mk_gen = 0;              // bumped whenever the marker buffers are reallocated
    int ngpart64 = 0;            // tile columns of the FP64 backward transform along y
    int ysh_avail = 0;           // n2m / 2 when the mirror form of the y transform is set up
    bool dph_mirror = false;     // dph holds the mirror form's result before its pair butterfly
    bool ymirror_on = true;      // the mirror form selected (used while the transform is along y in FP64)
    struct fribo_batch_ctx* batch = nullptr;   // batch this instance belongs to (shares its stream)
    int dev = 0;                 // device of the handle
    double* scratch = nullptr;   // device address of the static loads scratch on that device
};

template <class T>
static int dalloc(fribo_ctx* c, T** p, size_t n, std::vector<void*>* list = nullptr)
{
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
    (list ? list : &c->allocs)->push_back(q);
    *p = (T*)q;
    return FRIBO_OK;
}

// ---------------------------------------------------------------------------
// Grid metrics, restated from the reference set-up (host, once).
namespace {
enum { A_RC, A_RM, A_ZZ, A_ZM, A_G2RC, A_G2RM, A_G3RC, A_G3RM, A_UDX2C, A_UDX2M, A_UDX3C, A_UDX3M,
       A_AM2J, A_AC2J, A_AP2J, A_AM3J, A_AC3J, A_AP3J, A_AM3CK, A_AC3CK, A_AP3CK, A_AM3SK, A_AC3SK, A_AP3SK,
       A_AMPHJ, A_ACPHJ, A_APPHJ, A_AMPHK, A_ACPHK, A_APPHK, A_WR, A_COUNT };
const char* kArrNames[A_COUNT] = {"rc", "rm", "zz", "zm", "g2rc", "g2rm", "g3rc", "g3rm", "udx2c", "udx2m", "udx3c",
                                  "udx3m", "am2j", "ac2j", "ap2j", "am3j", "ac3j", "ap3j", "am3ck", "ac3ck", "ap3ck",
                                  "am3sk", "ac3sk", "ap3sk", "amphj", "acphj", "apphj", "amphk", "acphk", "apphk", "wr"};
bool is_zdir(int a)
{
    return a == A_ZZ || a == A_ZM || a == A_G3RC || a == A_G3RM || a == A_UDX3C || a == A_UDX3M || a == A_AM3CK ||
           a == A_AC3CK || a == A_AP3CK || a == A_AM3SK || a == A_AC3SK || a == A_AP3SK || a == A_AMPHK ||
           a == A_ACPHK || a == A_APPHK || a == A_WR;
}
}  // namespace

static void host_metrics(fribo_ctx* c, const double* rc_in, const double* zz_in, std::vector<int>& jmv,
                         std::vector<int>& jpv, std::vector<int>& kmv, std::vector<int>& kpv)
{
    const int n2 = c->p.n2, n3 = c->p.n3, n2m = n2 - 1, n3m = n3 - 1;
    auto& H = c->h1;
    for (int a = 0; a < A_COUNT; ++a) H[a].assign((is_zdir(a) ? n3 : n2) + 3, 0.0);
    const double dx2 = c->F.dx2, dx3 = c->F.dx3, dx2q = c->F.dx2q, dx3q = c->F.dx3q;
    auto& rc = H[A_RC]; auto& rm = H[A_RM]; auto& zz = H[A_ZZ]; auto& zm = H[A_ZM];
    auto& g2rc = H[A_G2RC]; auto& g2rm = H[A_G2RM]; auto& g3rc = H[A_G3RC]; auto& g3rm = H[A_G3RM];
    for (int j = 1; j <= n2; ++j) rc[j] = rc_in[j - 1];
    for (int k = 1; k <= n3; ++k) zz[k] = zz_in[k - 1];
    // indic.F90:99-125
    jmv.assign(n2 + 3, 0); jpv.assign(n2 + 3, 0); kmv.assign(n3 + 3, 0); kpv.assign(n3 + 3, 0);
    std::vector<int> jmc(n2 + 3, 0), jpc(n2 + 3, 0), kmc(n3 + 3, 0), kpc(n3 + 3, 0);
    for (int kc = 1; kc <= n3m; ++kc) {
        kmv[kc] = kc - 1; kpv[kc] = kc + 1;
        if (kc == 1) kmv[kc] = kc;
        if (kc == n3m) kpv[kc] = kc;
    }
    for (int kc = 1; kc <= n3m; ++kc) { kpc[kc] = kpv[kc] - kc; kmc[kc] = kc - kmv[kc]; }
    for (int jc = 1; jc <= n2; ++jc) {
        jmv[jc] = jc - 1; jpv[jc] = jc + 1;
        if (jc == 1) jmv[jc] = jc;
        if (jc == n2m) jpv[jc] = jc;
        if (jc == n2) jpv[jc] = jc;
    }
    for (int jc = 1; jc <= n2m; ++jc) { jpc[jc] = jpv[jc] - jc; jmc[jc] = jc - jmv[jc]; }
    // cordin.F90:106-123, 197-215
    for (int j = 1; j <= n2m; ++j) { rm[j] = (rc[j] + rc[j + 1]) * 0.5; g2rm[j] = (rc[j + 1] - rc[j]) * dx2; }
    for (int j = 2; j <= n2m; ++j) g2rc[j] = (rc[j + 1] - rc[j - 1]) * dx2 * 0.5;
    g2rc[1] = (rc[2] - rc[1]) * dx2;
    g2rc[n2] = (rc[n2] - rc[n2m]) * dx2;
    rm[n2] = rc[n2];
    for (int j = 1; j <= n2m; ++j) { H[A_UDX2M][j] = dx2 / g2rm[j]; H[A_UDX2C][j] = dx2 / g2rc[j]; }
    H[A_UDX2C][n2] = dx2 / g2rc[n2];
    for (int k = 1; k <= n3m; ++k) { zm[k] = (zz[k] + zz[k + 1]) * 0.5; g3rm[k] = (zz[k + 1] - zz[k]) * dx3; }
    for (int k = 2; k <= n3m; ++k) g3rc[k] = (zz[k + 1] - zz[k - 1]) * dx3 * 0.5;
    g3rc[1] = (zz[2] - zz[1]) * dx3;
    g3rc[n3] = (zz[n3] - zz[n3m]) * dx3;
    zm[n3] = zz[n3];
    for (int k = 1; k <= n3m; ++k) { H[A_UDX3M][k] = dx3 / g3rm[k]; H[A_UDX3C][k] = dx3 / g3rc[k]; }
    H[A_UDX3C][n3] = dx3 / g3rc[n3];
    // MLS_SD_ind reads one node past its grid copy at the last cell
    // (MLSint.F90:89,93); the nearest defined value is used.
    rm[n2 + 1] = rc[n2];
    zm[n3 + 1] = zz[n3];
    // coetar.F90:18-141
    const int inwn = c->p.inslwn ? 1 : 0, inws = c->p.inslws ? 1 : 0;
    auto& am2j = H[A_AM2J]; auto& ac2j = H[A_AC2J]; auto& ap2j = H[A_AP2J];
    auto& am3j = H[A_AM3J]; auto& ac3j = H[A_AC3J]; auto& ap3j = H[A_AP3J];
    am2j[1] = 0.0; ap2j[1] = 0.0; ac2j[1] = 1.0; am2j[n2] = 0.0; ap2j[n2] = 0.0; ac2j[n2] = 1.0;
    for (int jc = 2; jc <= n2m; ++jc) {
        int jm = jc - 1;
        double a22 = dx2q / g2rc[jc], a22p = 1.0 / g2rm[jc], a22m = 1.0 / g2rm[jm];
        ap2j[jc] = a22 * a22p; am2j[jc] = a22 * a22m; ac2j[jc] = -(a22 * a22p + a22 * a22m);
    }
    for (int jc = 2; jc <= n2m - 1; ++jc) {
        int jp = jpv[jc];
        double a22 = dx2q / g2rm[jc], a22p = a22 / g2rc[jp], a22m = a22 / g2rc[jc];
        ap3j[jc] = a22p; am3j[jc] = a22m; ac3j[jc] = -(a22p + a22m);
    }
    {
        int jc = n2m, jp = jc + 1;
        double a22 = dx2q / g2rm[jc], a22p = +a22 / g2rc[jp], a22m = +a22 / g2rc[jc];
        am3j[jc] = a22m; ap3j[jc] = 2.0 * inwn * a22p; ac3j[jc] = -(a22m + inwn * a22p * 2.0);
        jc = 1; jp = jc + 1;
        a22 = dx2q / g2rm[jc]; a22p = +a22 / g2rc[jp]; a22m = +a22 / g2rc[jc];
        ap3j[jc] = a22p; am3j[jc] = 2.0 * a22m * inws; ac3j[jc] = -(a22p + 2.0 * inws * a22m);
    }
    auto& am3ck = H[A_AM3CK]; auto& ac3ck = H[A_AC3CK]; auto& ap3ck = H[A_AP3CK];
    auto& am3sk = H[A_AM3SK]; auto& ac3sk = H[A_AC3SK]; auto& ap3sk = H[A_AP3SK];
    am3ck[1] = 0.0; ap3ck[1] = 0.0; ac3ck[1] = 1.0; am3ck[n3] = 0.0; ap3ck[n3] = 0.0; ac3ck[n3] = 1.0;
    for (int kc = 2; kc <= n3m; ++kc) {
        int km = kc - 1;
        double a33 = dx3q / g3rc[kc], a33p = 1.0 / g3rm[kc], a33m = 1.0 / g3rm[km];
        ap3ck[kc] = a33 * a33p; am3ck[kc] = a33 * a33m; ac3ck[kc] = -(ap3ck[kc] + am3ck[kc]);
    }
    for (int kc = 2; kc <= n3m - 1; ++kc) {
        int kp = kc + 1;
        double a33 = dx3q / g3rm[kc], a33p = +a33 / g3rc[kp], a33m = +a33 / g3rc[kc];
        ap3sk[kc] = a33p; am3sk[kc] = a33m; ac3sk[kc] = -(ap3sk[kc] + am3sk[kc]);
    }
    {
        int kc = 1, kp = kc + 1;
        double a33 = dx3q / g3rm[kc], a33p = +a33 / g3rc[kp], a33m = +a33 / g3rc[kc];
        ap3sk[kc] = a33p; am3sk[kc] = 0.0; ac3sk[kc] = -(a33p + a33m * 2.0);
        kc = n3m; kp = kc + 1;
        a33 = dx3q / g3rm[kc]; a33p = +a33 / g3rc[kp]; a33m = +a33 / g3rc[kc];
        am3sk[kc] = a33m; ap3sk[kc] = 0.0; ac3sk[kc] = -(a33m + a33p * 2.0);
    }
    // phini.F90:121-142 (tridiag_matrices)
    for (int jc = 1; jc <= n2m; ++jc) {
        int jp = jpv[jc];
        double a22icc = jmc[jc] * dx2q / g2rc[jc], a22icp = jpc[jc] * dx2q / g2rc[jp], ug = 1.0 / g2rm[jc];
        H[A_AMPHJ][jc] = a22icc * ug; H[A_APPHJ][jc] = a22icp * ug;
        H[A_ACPHJ][jc] = -(H[A_AMPHJ][jc] + H[A_APPHJ][jc]);
    }
    for (int kc = 1; kc <= n3m; ++kc) {
        int kp = kpv[kc];
        double a33icc = kmc[kc] * dx3q / g3rc[kc], a33icp = kpc[kc] * dx3q / g3rc[kp], ug = 1.0 / g3rm[kc];
        H[A_AMPHK][kc] = a33icc * ug; H[A_APPHK][kc] = a33icp * ug;
        H[A_ACPHK][kc] = -(H[A_AMPHK][kc] + H[A_APPHK][kc]);
    }
}

static int upload1d(fribo_ctx* c, double** dst, const std::vector<double>& h);

// Partition rows [0,n) into nch = min(64, n/2) chunks starting at floor(p*n/nch).
static std::vector<int> chunk_starts(int n, int* nch_out, int* maxm_out)
{
    int nch = n / 2;
    if (nch > FRIBO_KMAXCH) nch = FRIBO_KMAXCH;
    if (nch < 1) nch = 1;
    std::vector<int> cs(nch + 1);
    int maxm = 0;
    for (int p = 0; p <= nch; ++p) cs[p] = (int)(((long)p * n) / nch);
    for (int p = 0; p < nch; ++p) maxm = std::max(maxm, cs[p + 1] - cs[p]);
    *nch_out = nch;
    *maxm_out = maxm;
    return cs;
}

static int setup_ksys(fribo_ctx* c, KSys& K, int n, int kofs, int src, int inter)
{
    K.n = n;
    K.kofs = kofs;
    K.src = src;
    K.inter = inter;
    std::vector<int> cs = chunk_starts(n, &K.nch, &K.maxm);
    if (K.maxm - 1 > (inter ? 64 : 128))
        return set_err(FRIBO_ERR_LIMIT, "tridiagonal lines too long for this build");
    int rc;
    if ((rc = dalloc(c, &K.cs, K.nch + 1))) return rc;
    HIPCHK(hipMemcpy(K.cs, cs.data(), (K.nch + 1) * sizeof(int), hipMemcpyHostToDevice));
    const size_t len = inter ? (size_t)K.maxm * 64 : (size_t)n;
    double** arrs[] = {&K.a, &K.c, &K.S, &K.cp, &K.inv, &K.vp, &K.g};
    for (auto a : arrs)
        if ((rc = dalloc(c, a, len))) return rc;
    if ((rc = dalloc(c, &K.wl, 64))) return rc;
    if ((rc = dalloc(c, &K.k1, 6 * 64))) return rc;
    if ((rc = dalloc(c, &K.k2, 6 * 64))) return rc;
    if ((rc = dalloc(c, &K.ib, 64))) return rc;
    if ((rc = dalloc(c, &K.bc, 2))) return rc;
    return FRIBO_OK;
}

// Poisson j-lines: fixed partition and lane-interleaved amphj/acphj/apphj.
static int setup_psys(fribo_ctx* c, PSys& Q, int n, const std::vector<double>& am, const std::vector<double>& ac,
                      const std::vector<double>& ap)
{
    Q.n = n;
    std::vector<int> cs = chunk_starts(n, &Q.nch, &Q.maxm);
    if (Q.maxm - 1 > 64) return set_err(FRIBO_ERR_LIMIT, "Poisson lines too long for this build");
    int rc;
    if ((rc = dalloc(c, &Q.cs, Q.nch + 1))) return rc;
    HIPCHK(hipMemcpy(Q.cs, cs.data(), (Q.nch + 1) * sizeof(int), hipMemcpyHostToDevice));
    const size_t len = (size_t)Q.maxm * 64;
    std::vector<double> ham(len, 0.0), hac(len, 1.0), hap(len, 0.0);
    for (int p = 0; p < Q.nch; ++p)
        for (int r = cs[p]; r < cs[p + 1]; ++r) {
            const size_t ix = (size_t)(r - cs[p]) * 64 + p;
            const int j = r + 1;
            ham[ix] = am[j];
            hac[ix] = ac[j];
            hap[ix] = ap[j];
        }
    if ((rc = upload1d(c, &Q.am, ham))) return rc;
    if ((rc = upload1d(c, &Q.ac, hac))) return rc;
    if ((rc = upload1d(c, &Q.ap, hap))) return rc;
    return FRIBO_OK;
}

static int upload1d(fribo_ctx* c, double** dst, const std::vector<double>& h)
{
    int rc = dalloc(c, dst, h.size());
    if (rc) return rc;
    HIPCHK(hipMemcpy(*dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return FRIBO_OK;
}

static int upload1i(fribo_ctx* c, int** dst, const std::vector<int>& h)
{
    int rc = dalloc(c, dst, h.size());
    if (rc) return rc;
    HIPCHK(hipMemcpy(*dst, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    return FRIBO_OK;
}

// Gauge partials the backward transform along y writes: one per tile column
// (two in the mirror form), or those of the FP32 transform.
static int ngpart_for(const fribo_ctx* c)
{
    if (c->F.refine >= 0) return gemm32_tiles_n(c->F.n3m);
    return c->F.ysh > 0 ? 2 * c->ngpart64 : c->ngpart64;
}

// F.ysh from the options: n2m / 2 while the mirror form runs (selected,
// available, transform along y in FP64), else 0; the kernels read only F.ysh.
static void apply_ysh(fribo_ctx* c)
{
    FluidDev& F = c->F;
    F.ysh = (c->ymirror_on && c->ysh_avail > 0 && F.taxis == 1 && F.refine < 0) ? c->ysh_avail : 0;
    if (F.taxis == 1) F.ngpart = ngpart_for(c);
}

// Eigen-system of the y pressure operator (fribo_eig.hip) and the gauge
// weights zg = zmmt(null_mode, :); leaves taxis = 1 when it succeeds and the
// y operator has a single null mode, otherwise keeps the transform along z.
static int setup_ytransform(fribo_ctx* c, const double* zmmt, int ldz)
{
    FluidDev& F = c->F;
    const int n = F.n2m;
    F.ldy = (n + 15) / 16 * 16;
    std::vector<double> Y((size_t)F.ldy * n, 0.0), Yi((size_t)F.ldy * n, 0.0), wy(n);
    if (ytransform_build(n, c->h1[A_AMPHJ].data() + 1, c->h1[A_ACPHJ].data() + 1, c->h1[A_APPHJ].data() + 1, F.ldy,
                         Y.data(), Yi.data(), wy.data()) != 0)
        return FRIBO_OK;
    double wmax = 0.0, wmin = 1e300;
    int imin = 0;
    for (int m = 0; m < n; ++m) {
        const double a = fabs(wy[m]);
        wmax = std::max(wmax, a);
        if (a < wmin) { wmin = a; imin = m; }
    }
    if (!(wmin <= 1e-9 * wmax)) return FRIBO_OK;
    std::vector<double> zg(F.n3m);
    double zs = 0.0;
    for (int k = 0; k < F.n3m; ++k) {
        zg[k] = zmmt[(size_t)(F.null_mode - 1) + (size_t)k * ldz];
        zs = zs + zg[k];
    }
    if (!(fabs(zs) > 0.0)) return FRIBO_OK;
    int e = 0;
    if (!e) e = dalloc(c, &F.ymm, (size_t)F.ldy * n);
    if (!e) e = dalloc(c, &F.ymmi, (size_t)F.ldy * n);
    if (!e) e = dalloc(c, &F.wy, n);
    if (!e) e = dalloc(c, &F.zg, F.n3m);
    if (!e) e = dalloc(c, &F.ysum, (size_t)7 * n * FRIBO_KMAXCH);
    c->ngpart64 = gemm_nn_tiles_n(n, F.n3m, n, F.ymm, F.ldy, F.hmn, F.P);
    // room for the gauge partials of any tile shape (a batch may use
    // 32-column tiles) and of the mirror form (two per tile column)
    if (!e) e = dalloc(c, &F.gpart, 2 * std::max(c->ngpart64, (F.n3m + 31) / 32));
    if (e) return e;
    HIPCHK(hipMemcpy(F.ymm, Y.data(), Y.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.ymmi, Yi.data(), Yi.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.wy, wy.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.zg, zg.data(), F.n3m * sizeof(double), hipMemcpyHostToDevice));
    F.ynull = imin + 1;
    F.zginv = 1.0 / zs;
    F.taxis = 1;
    c->yavail = true;
    // the mirror form, when the grid is symmetric about the channel centre
    // (coefficients equal to their mirror image to 1e-10 of the largest)
    if (ytransform_mirror(n, c->h1[A_AMPHJ].data() + 1, c->h1[A_ACPHJ].data() + 1, c->h1[A_APPHJ].data() + 1, F.ldy,
                          Y.data(), Yi.data(), wy.data(), 1e-10) == 0) {
        double smin = 1e300;
        int is = 0;
        for (int m = 0; m < n; ++m)
            if (fabs(wy[m]) < smin) { smin = fabs(wy[m]); is = m; }
        if (smin <= 1e-9 * wmax) {
            if (!e) e = dalloc(c, &F.ysm, (size_t)F.ldy * n);
            if (!e) e = dalloc(c, &F.ysmi, (size_t)F.ldy * n);
            if (!e) e = dalloc(c, &F.wys, n);
            if (e) return e;
            HIPCHK(hipMemcpy(F.ysm, Y.data(), Y.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(F.ysmi, Yi.data(), Yi.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(F.wys, wy.data(), n * sizeof(double), hipMemcpyHostToDevice));
            F.ynulls = is + 1;
            c->ysh_avail = n / 2;
        }
    }
    apply_ysh(c);
    return FRIBO_OK;
}

// Static device scratch of fribo_loads_transfer_static: the probe inputs
// lmnor, lmacc (2 nlm each), the outputs press (nlm) and tau (2 nlm) and the
// per-marker work records (FRIBO_LWORK nlm) for up to
// FRIBO_STATIC_LOADS_MAX_NLM markers, 1.8 MB held on every device the
// library runs on until the process ends.  Host code never passes the
// array's name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress returns, once per device, kept in g_scratch_addr.
static __device__ __align__(256) double scratch[(size_t)FRIBO_STATIC_LOADS_MAX_NLM * (7 + FRIBO_LWORK)];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_addr[kMaxDevices] = {nullptr};

// Device address of `scratch` on device `dev`, which must be current.
static int scratch_address(int dev, double** out)
{
    if (dev < 0 || dev >= kMaxDevices) return set_err(FRIBO_ERR_LIMIT, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !p)
            return set_err(FRIBO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(he));
        g_scratch_addr[dev] = (double*)p;
    }
    *out = g_scratch_addr[dev];
    return FRIBO_OK;
}

extern "C" int fribo_init(const fribo_params* p, const double* rc, const double* zz, const double* zmm,
                          const double* zmmt, int ldz, const double* wr, fribo_handle* out)
{
    if (!p || !rc || !zz || !out) return set_err(FRIBO_ERR_ARG, "null argument");
    if (!(zmm && zmmt && wr) && (zmm || zmmt || wr))
        return set_err(FRIBO_ERR_ARG, "zmm, zmmt and wr are given together or not at all");
    // no phini inputs: the library decomposes the z operator itself
    std::vector<double> own_zmm, own_zmmt, own_wr;
    if (!zmm) {
        const int n3m = p->n3 - 1;
        if (n3m < 2) return set_err(FRIBO_ERR_ARG, "n3 too small");
        own_zmm.assign((size_t)n3m * n3m, 0.0);
        own_zmmt.assign((size_t)n3m * n3m, 0.0);
        own_wr.assign(n3m, 0.0);
        if (phini_sym(p->n3, zz, p->alx3, n3m, own_zmm.data(), own_zmmt.data(), own_wr.data()) != 0)
            return set_err(FRIBO_ERR_LIMIT, "eigen-decomposition of the z pressure operator failed");
        zmm = own_zmm.data();
        zmmt = own_zmmt.data();
        wr = own_wr.data();
        ldz = n3m;
    }
    if (p->n2 < 5 || p->n3 < 5) return set_err(FRIBO_ERR_ARG, "grid too small (n2, n3 >= 5)");
    if (p->nsst != 1 && p->nsst != 3) return set_err(FRIBO_ERR_ARG, "nsst must be 1 or 3");
    if (ldz < p->n3 - 1) return set_err(FRIBO_ERR_ARG, "ldz < n3m");
    if (p->wfun < 1 || p->wfun > 4) return set_err(FRIBO_ERR_ARG, "wfun must be 1..4");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    double* scratch_dev = nullptr;
    if (int se = scratch_address(dev, &scratch_dev)) return se;
    fribo_ctx* c = new fribo_ctx();
    c->p = *p;
    c->dev = dev;
    c->scratch = scratch_dev;
    c->lds_limit = prop.sharedMemPerBlock;
    FluidDev& F = c->F;
    memset(&F, 0, sizeof(F));
    F.refine = -1;
    F.n2 = p->n2; F.n3 = p->n3; F.n2m = p->n2 - 1; F.n3m = p->n3 - 1;
    F.lds_max = (int)c->lds_limit;
    F.P = (F.n2 + 15) / 16 * 16;
    F.nk = F.n3m + 6;
    F.nsst = p->nsst; F.idtv = p->idtv; F.ibact = p->ibact; F.wfun = p->wfun;
    F.inslws = p->inslws; F.inslwn = p->inslwn;
    F.alx3 = p->alx3; F.rext2 = p->rext2; F.cou = p->cou; F.q3uw = p->q3uw; F.q3lw = p->q3lw;
    F.dtmax = p->dtmax; F.cflmax = p->cflmax; F.cfllim = p->cfllim; F.radiusSD = p->radiusSD;
    // fluid_pre.F90:33-70
    F.dx2 = 1.0 / (p->rext2 / (double)F.n2m);
    F.dx3 = 1.0 / (p->alx3 / (double)F.n3m);
    F.dx2q = F.dx2 * F.dx2;
    F.dx3q = F.dx3 * F.dx3;
    F.ren = 1.0 / p->visc;
    F.visc = p->visc;
    if (p->nsst > 1) {
        F.gam[1] = 8.0 / 15.0; F.gam[2] = 5.0 / 12.0; F.gam[3] = 3.0 / 4.0;
        F.rom[1] = 0.0; F.rom[2] = -17.0 / 60.0; F.rom[3] = -5.0 / 12.0;
    } else {
        F.gam[1] = 1.5; F.rom[1] = -0.50;
    }
    for (int n = 1; n <= p->nsst; ++n) F.alm[n] = F.gam[n] + F.rom[n];

    std::vector<int> jmv, jpv, kmv, kpv;
    host_metrics(c, rc, zz, jmv, jpv, kmv, kpv);
    for (int k = 1; k <= F.n3m; ++k) c->h1[A_WR][k] = wr[k - 1];
    // null mode of the Neumann z operator
    {
        double wmax = 0.0, wmin = 1e300;
        int imin = 0;
        for (int k = 1; k <= F.n3m; ++k) {
            double a = fabs(wr[k - 1]);
            wmax = std::max(wmax, a);
            if (a < wmin) { wmin = a; imin = k; }
        }
        F.null_mode = (wmin <= 1e-9 * wmax) ? imin : 0;
    }
    int e = 0;
    double** d1[A_COUNT] = {&F.rc, &F.rm, &F.zz, &F.zm, &F.g2rc, &F.g2rm, &F.g3rc, &F.g3rm, &F.udx2c, &F.udx2m,
                            &F.udx3c, &F.udx3m, &F.am2j, &F.ac2j, &F.ap2j, &F.am3j, &F.ac3j, &F.ap3j, &F.am3ck,
                            &F.ac3ck, &F.ap3ck, &F.am3sk, &F.ac3sk, &F.ap3sk, &F.amphj, &F.acphj, &F.apphj,
                            &F.amphk, &F.acphk, &F.apphk, &F.wr};
    for (int a = 0; a < A_COUNT && !e; ++a) e = upload1d(c, d1[a], c->h1[a]);
    if (!e) e = upload1i(c, &F.jmv, jmv);
    if (!e) e = upload1i(c, &F.jpv, jpv);
    if (!e) e = upload1i(c, &F.kmv, kmv);
    if (!e) e = upload1i(c, &F.kpv, kpv);
    const size_t fsz = (size_t)F.P * F.nk;
    double** flds[] = {&F.q2, &F.q3, &F.pr, &F.dph, &F.ru2, &F.ru3, &F.rhs2, &F.rhs3, &F.sA, &F.sB, &F.hmn};
    for (auto f : flds)
        if (!e) e = dalloc(c, f, fsz);
    double** bcs[] = {&F.qb2s, &F.qb3s, &F.qb2n, &F.qb3n, &F.dqb2s, &F.dqb3s, &F.dqb2n, &F.dqb3n, &F.dq2x2o, &F.dq3x2o};
    for (auto b : bcs)
        if (!e) e = dalloc(c, b, F.n2 + 3);
    // Poisson matrices, packed with a 16-aligned leading dimension
    F.ldz = (F.n3m + 15) / 16 * 16;
    if (!e) e = dalloc(c, &F.zmm, (size_t)F.ldz * F.n3m);
    if (!e) e = dalloc(c, &F.zmmt, (size_t)F.ldz * F.n3m);
    if (!e) {
        hipError_t he = hipMemcpy2D(F.zmm, F.ldz * sizeof(double), zmm, ldz * sizeof(double), F.n3m * sizeof(double),
                                    F.n3m, hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy2D(F.zmmt, F.ldz * sizeof(double), zmmt, ldz * sizeof(double), F.n3m * sizeof(double), F.n3m,
                             hipMemcpyHostToDevice);
        if (he != hipSuccess) e = set_err(FRIBO_ERR_HIP, std::string("hipMemcpy2D (zmm, zmmt): ") + hipGetErrorString(he));
    }
    F.ncfl = fribo_ncfl(F.n2m, F.n3m);
    if (!e) e = dalloc(c, &F.cfl_part, 2 * (size_t)F.ncfl);
    F.npr_part = fribo_npr_part(F.n2m, F.n3m);
    if (!e) e = dalloc(c, &F.pr_part, F.npr_part);
    if (!e) e = setup_ksys(c, F.ks[0], F.n3m, 1, 0, 0);
    if (!e) e = setup_ksys(c, F.ks[1], F.n3m - 1, 2, 0, 0);
    if (!e) e = setup_ksys(c, F.js[0], F.n2 - 2, 2, 1, 1);
    if (!e) e = setup_ksys(c, F.js[1], F.n2m, 1, 2, 1);
    F.kform = ksolve_fused_fits(F) ? 1 : 0;
    if (!e) e = setup_psys(c, F.ps, F.n2m, c->h1[A_AMPHJ], c->h1[A_ACPHJ], c->h1[A_APPHJ]);
    if (!e) {
        F.jstride = jsolve_stride(F);
        if (F.n2m > 256 * 16)
            e = set_err(FRIBO_ERR_LIMIT, "n2 too large for the CFL kernel of this build (n2-1 <= 4096)");
        else if (jsolve_lds_bytes(F) > c->lds_limit || sysprep_lds_bytes(F) + 17000 > c->lds_limit)
            e = set_err(FRIBO_ERR_LIMIT, "grid too large for the LDS staging of this build");
    }
    if (!e) e = dalloc(c, &F.kYs, (size_t)2 * FRIBO_KMAXCH * F.n2m);
    if (!e) e = dalloc(c, &F.kYe, (size_t)2 * FRIBO_KMAXCH * F.n2m);
    if (!e) e = dalloc(c, &F.kDe, (size_t)2 * FRIBO_KMAXCH * F.n2m);
    if (!e) e = dalloc(c, &F.kX, (size_t)2 * FRIBO_KMAXCH * F.n2m);
    if (!e) e = dalloc(c, &F.S, 1);
    if (!e) e = dalloc(c, &c->nb, 2);
    if (!e) e = dalloc(c, &c->dpart, 8 * 256 + 16);
    {
        // utils_proc.F90:600-602: first node nearest to z = 18.5
        int indz = 1;
        double best = fabs(zz[0] - 18.5);
        for (int k = 2; k <= F.n3; ++k)
            if (fabs(zz[k - 1] - 18.5) < best) { best = fabs(zz[k - 1] - 18.5); indz = k; }
        c->vort_k = (indz <= F.n3m) ? indz : 0;
    }
    {
        double dymin = 1.0e+6, dzmin = 1.0e+6;
        for (int i = 2; i <= F.n2; ++i) dymin = std::min(dymin, rc[i - 1] - rc[i - 2]);
        for (int i = 2; i <= F.n3; ++i) dzmin = std::min(dzmin, zz[i - 1] - zz[i - 2]);
        c->dmin = std::min(dzmin, dymin);
    }
    // the Poisson transform along y costs 2*n3m*n2m^2 flops against
    // 2*n2m*n3m^2 along z: set it up when y is the shorter axis
    if (!e && F.n2m < F.n3m && F.null_mode > 0) e = setup_ytransform(c, zmmt, ldz);
    if (!e) {
        hipError_t he = hipStreamCreate(&c->st);
        if (he != hipSuccess) e = set_err(FRIBO_ERR_HIP, hipGetErrorString(he));
    }
    if (e) {
        for (void* q : c->allocs) hipFree(q);
        delete c;
        return e;
    }
    memset(&c->hs, 0, sizeof(c->hs));
    {
        hipError_t he = hipHostMalloc((void**)&c->in, sizeof(fribo_ctx::StepIn), hipHostMallocDefault);
        if (he == hipSuccess) he = hipHostMalloc((void**)&c->hsp, sizeof(DevScalars), hipHostMallocDefault);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_rb, hipEventDisableTiming);
        if (he != hipSuccess) {
            for (void* q : c->allocs) hipFree(q);
            if (c->in) hipHostFree(c->in);
            if (c->hsp) hipHostFree(c->hsp);
            if (c->ev_rb) hipEventDestroy(c->ev_rb);
            hipStreamDestroy(c->st);
            delete c;
            return set_err(FRIBO_ERR_HIP, hipGetErrorString(he));
        }
    }
    *out = c;
    return FRIBO_OK;
}

extern "C" int fribo_finalize(fribo_handle h)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    if (h->batch) return set_err(FRIBO_ERR_ARG, "instance belongs to a batch: destroy the batch first");
    hipStreamSynchronize(h->st);
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    if (h->in) hipHostFree(h->in);
    if (h->hsp) hipHostFree(h->hsp);
    if (h->ev_rb) hipEventDestroy(h->ev_rb);
    for (auto ev : h->evpool) hipEventDestroy(ev);
    for (void* q : h->mk_allocs) hipFree(q);
    if (h->mk_stage) hipHostFree(h->mk_stage);
    if (h->ev_mk) hipEventDestroy(h->ev_mk);
    for (void* q : h->allocs) hipFree(q);
    hipStreamDestroy(h->st);
    delete h;
    return FRIBO_OK;
}

// ---------------------------------------------------------------------------
// State transfer (pitched copies between the reference layout and the device)
static int put_field(fribo_ctx* c, double* dev, const double* host, int kfirst, int nplanes)
{
    const FluidDev& F = c->F;
    HIPCHK(hipMemcpy2DAsync(dev + fidx(F, 1, kfirst), F.P * sizeof(double), host, F.n2 * sizeof(double),
                            F.n2 * sizeof(double), nplanes, hipMemcpyHostToDevice, c->st));
    return FRIBO_OK;
}

static int get_field(fribo_ctx* c, double* host, const double* dev, int kfirst, int nplanes)
{
    const FluidDev& F = c->F;
    HIPCHK(hipMemcpy2DAsync(host, F.n2 * sizeof(double), dev + fidx(F, 1, kfirst), F.P * sizeof(double),
                            F.n2 * sizeof(double), nplanes, hipMemcpyDeviceToHost, c->st));
    return FRIBO_OK;
}

static void enqueue_cfl(fribo_ctx* c);

// next_cfl: after the copy, enqueue the next step's k_cfl and wait for the
// copy only.
static int read_scalars(fribo_ctx* c, bool next_cfl = false)
{
    HIPCHK(hipMemcpyAsync(c->hsp, c->F.S, sizeof(DevScalars), hipMemcpyDeviceToHost, c->st));
    if (next_cfl) {
        HIPCHK(hipEventRecord(c->ev_rb, c->st));
        enqueue_cfl(c);
        HIPCHK(hipEventSynchronize(c->ev_rb));
    } else {
        HIPCHK(hipStreamSynchronize(c->st));
    }
    c->hs = *c->hsp;
    c->cfl_ready = next_cfl && c->hs.err == 0;
    return FRIBO_OK;
}

extern "C" int fribo_set_state(fribo_handle h, const double* q2, const double* q3, const double* pr,
                               const double* ru2, const double* ru3, const double* qb2s, const double* qb3s,
                               const double* qb2n, const double* qb3n, const double* dq2x2o, const double* dq3x2o)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    h->cfl_ready = false;
    FluidDev& F = h->F;
    int e = 0;
    if (q2 && !e) e = put_field(h, F.q2, q2, -2, F.nk);
    if (q3 && !e) e = put_field(h, F.q3, q3, -2, F.nk);
    if (pr && !e) e = put_field(h, F.pr, pr, -2, F.nk);
    if (ru2 && !e) e = put_field(h, F.ru2, ru2, 1, F.n3m);
    if (ru3 && !e) e = put_field(h, F.ru3, ru3, 1, F.n3m);
    const double* b[] = {qb2s, qb3s, qb2n, qb3n, dq2x2o, dq3x2o};
    double* d[] = {F.qb2s, F.qb3s, F.qb2n, F.qb3n, F.dq2x2o, F.dq3x2o};
    for (int i = 0; i < 6 && !e; ++i)
        if (b[i]) {
            hipError_t he = hipMemcpyAsync(d[i] + 1, b[i], F.n2 * sizeof(double), hipMemcpyHostToDevice, h->st);
            if (he != hipSuccess) e = set_err(FRIBO_ERR_HIP, hipGetErrorString(he));
        }
    if (e) return e;
    if (pr) {
        // the uploaded pressure is already mean-free: nothing pending
        if ((e = read_scalars(h))) return e;
        h->hs.pend = 0.0;
        h->hs.goff = 0.0;
        static_assert(offsetof(DevScalars, goff) == offsetof(DevScalars, pend) + sizeof(double), "layout");
        HIPCHK(hipMemcpyAsync(&F.S->pend, &h->hs.pend, 2 * sizeof(double), hipMemcpyHostToDevice, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return FRIBO_OK;
}

extern "C" int fribo_get_state(fribo_handle h, double* q2, double* q3, double* pr, double* ru2, double* ru3,
                               double* qb2s, double* qb3s, double* qb2n, double* qb3n, double* dq2x2o,
                               double* dq3x2o)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    FluidDev& F = h->F;
    int e = 0;
    if (q2 && !e) e = get_field(h, q2, F.q2, -2, F.nk);
    if (q3 && !e) e = get_field(h, q3, F.q3, -2, F.nk);
    if (pr && !e) e = get_field(h, pr, F.pr, -2, F.nk);
    if (ru2 && !e) e = get_field(h, ru2, F.ru2, 1, F.n3m);
    if (ru3 && !e) e = get_field(h, ru3, F.ru3, 1, F.n3m);
    double* b[] = {qb2s, qb3s, qb2n, qb3n, dq2x2o, dq3x2o};
    double* d[] = {F.qb2s, F.qb3s, F.qb2n, F.qb3n, F.dq2x2o, F.dq3x2o};
    for (int i = 0; i < 6 && !e; ++i)
        if (b[i]) {
            hipError_t he = hipMemcpyAsync(b[i], d[i] + 1, F.n2 * sizeof(double), hipMemcpyDeviceToHost, h->st);
            if (he != hipSuccess) e = set_err(FRIBO_ERR_HIP, hipGetErrorString(he));
        }
    if (e) return e;
    if ((e = read_scalars(h))) return e;
    if (pr) {
        // apply the pending means (updvp.F90:141): the last one to the
        // interior, their running sum to row n2 and the ghost planes
        const double pend = h->hs.pend, goff = h->hs.goff;
        for (int k = -2; k <= F.n3m + 3; ++k) {
            const bool ghost = (k < 1 || k > F.n3m);
            for (int j = 1; j <= F.n2; ++j) {
                double& v = pr[(size_t)(j - 1) + (size_t)(k + 2) * F.n2];
                v = v - ((ghost || j == F.n2) ? goff : pend);
            }
        }
    }
    return FRIBO_OK;
}

extern "C" int fribo_set_markers(fribo_handle h, int32_t nlm, const double* lmcc, const double* lmvel,
                                 const double* lmA)
{
    if (!h || nlm < 0 || (nlm > 0 && (!lmcc || !lmvel || !lmA))) return set_err(FRIBO_ERR_ARG, "bad marker args");
    FluidDev& F = h->F;
    if (nlm != h->nlm_alloc) {
        HIPCHK(hipStreamSynchronize(h->st));
        // the captured step holds the old marker buffers in its kernel
        // arguments: it is rebuilt on the next graph launch
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
        h->graph_nlm = -1;
        ++h->mk_gen;
        for (void* q : h->mk_allocs) hipFree(q);
        h->mk_allocs.clear();
        if (h->mk_stage) hipHostFree(h->mk_stage);
        h->mk_stage = nullptr;
        h->nlm_alloc = -1;   // no marker buffers until the allocation below succeeds
        F.nlm = 0;
        int maxkeys = 2048;
        while (maxkeys < nlm * FRIBO_DIMSD) maxkeys <<= 1;
        if (nlm >= (1 << 20) / FRIBO_DIMSD)
            return set_err(FRIBO_ERR_LIMIT, "too many markers (at most 116508)");
        F.maxkeys = maxkeys;
        auto* L = &h->mk_allocs;
        int e = 0;
        if (!e) e = dalloc(h, &F.lmcc, 5 * (size_t)nlm, L);
        if (!e) {
            F.lmvel = F.lmcc + 2 * (size_t)nlm;
            F.lmA = F.lmcc + 4 * (size_t)nlm;
        }
        if (!e) e = dalloc(h, &F.lmind, 4 * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.eulins, 4 * FRIBO_DIMSD * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.phi, 2 * FRIBO_DIMSD * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.phipred, 2 * FRIBO_DIMSD * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.cl, 2 * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.active, (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.contrib, 2 * FRIBO_DIMSD * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.keys, 2 * (size_t)maxkeys, L);
        if (!e) e = dalloc(h, &F.segval, 2 * (size_t)maxkeys, L);
        if (!e) e = dalloc(h, &F.lmnor, 2 * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.lmacc, 2 * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.press, (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.tau, 2 * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.lwork, FRIBO_LWORK * (size_t)nlm, L);
        if (!e) e = dalloc(h, &F.bcerr, 2 * (size_t)nlm, L);
        if (!e && nlm > 0) {
            hipError_t he = hipHostMalloc((void**)&h->mk_stage, 5 * (size_t)nlm * sizeof(double), hipHostMallocDefault);
            if (he == hipSuccess && !h->ev_mk) he = hipEventCreateWithFlags(&h->ev_mk, hipEventDisableTiming);
            if (he != hipSuccess) e = set_err(FRIBO_ERR_HIP, hipGetErrorString(he));
        }
        if (e) return e;
        h->nlm_alloc = nlm;
        F.nlm = nlm;
    }
    if (nlm > 0) {
        // the previous upload has left the staging buffer (normally long ago)
        HIPCHK(hipEventSynchronize(h->ev_mk));
        double* sbuf = h->mk_stage;
        memcpy(sbuf, lmcc, 2 * (size_t)nlm * sizeof(double));
        memcpy(sbuf + 2 * (size_t)nlm, lmvel, 2 * (size_t)nlm * sizeof(double));
        memcpy(sbuf + 4 * (size_t)nlm, lmA, (size_t)nlm * sizeof(double));
        HIPCHK(hipMemcpyAsync(F.lmcc, sbuf, 5 * (size_t)nlm * sizeof(double), hipMemcpyHostToDevice, h->st));
        HIPCHK(hipEventRecord(h->ev_mk, h->st));
        h->lmcc.assign(lmcc, lmcc + 2 * (size_t)nlm);
    }
    return FRIBO_OK;
}

// Windowed read-back: planes kfirst..klast (reference k indices) of the
// fields selected by mask (1 q2, 2 q3, 4 pr, 8 ru2, 16 ru3) into the host
// arrays of reference extents (q2, q3, pr: (1:n2, -2:n3m+3); ru2, ru3:
// (1:n2, 1:n3m), where the window is clipped to 1..n3m); the other planes of
// the host arrays are left as they are.  The pressure gets the pending
// means as in fribo_get_state.
extern "C" int fribo_fetch_fields(fribo_handle h, int32_t mask, int32_t kfirst, int32_t klast, double* q2,
                                  double* q3, double* pr, double* ru2, double* ru3)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    FluidDev& F = h->F;
    if (kfirst < -2 || klast > F.n3m + 3 || kfirst > klast) return set_err(FRIBO_ERR_ARG, "bad plane window");
    double* hf[5] = {q2, q3, pr, ru2, ru3};
    const double* df[5] = {F.q2, F.q3, F.pr, F.ru2, F.ru3};
    int e = 0;
    for (int f = 0; f < 5 && !e; ++f) {
        if (!((mask >> f) & 1)) continue;
        if (!hf[f]) return set_err(FRIBO_ERR_ARG, "null array for a selected field");
        const bool ru = f >= 3;
        const int a = ru ? std::max<int>(kfirst, 1) : kfirst, b = ru ? std::min<int>(klast, F.n3m) : klast;
        if (a > b) continue;
        const int base = ru ? 1 : -2;
        e = get_field(h, hf[f] + (size_t)(a - base) * F.n2, df[f], a, b - a + 1);
    }
    if (e) return e;
    if ((e = read_scalars(h))) return e;
    if ((mask & 4) && pr) {
        const double pend = h->hs.pend, goff = h->hs.goff;
        for (int k = kfirst; k <= klast; ++k) {
            const bool ghost = (k < 1 || k > F.n3m);
            for (int j = 1; j <= F.n2; ++j) {
                double& v = pr[(size_t)(j - 1) + (size_t)(k + 2) * F.n2];
                v = v - ((ghost || j == F.n2) ? goff : pend);
            }
        }
    }
    return FRIBO_OK;
}

// ---------------------------------------------------------------------------
// Step orchestration
static hipEvent_t next_event(fribo_ctx* c)
{
    if (c->evused == c->evpool.size()) {
        hipEvent_t ev;
        hipEventCreate(&ev);
        c->evpool.push_back(ev);
    }
    return c->evpool[c->evused++];
}

template <class Fn>
static void run(fribo_ctx* c, int cls, Fn fn)
{
    if (!c->profiling || !((c->prof_mask >> cls) & 1u)) {
        fn();
        return;
    }
    hipEvent_t a = next_event(c), b = next_event(c);
    hipEventRecord(a, c->st);
    fn();
    hipEventRecord(b, c->st);
    c->pending.push_back({cls, a, b});
}

// As run, for one launch that records its own start and end events (the
// dispatch's timestamps): timing puts no event packets between kernels.
template <class Fn>
static void run_ev(fribo_ctx* c, int cls, Fn fn)
{
    if (!c->profiling || !((c->prof_mask >> cls) & 1u)) {
        fn(nullptr, nullptr);
        return;
    }
    hipEvent_t a = next_event(c), b = next_event(c);
    fn(a, b);
    c->pending.push_back({cls, a, b});
}

static void collect_profile(fribo_ctx* c)
{
    for (auto& r : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->prof_ms[r.cls] += ms;
            c->prof_n[r.cls] += 1;
        }
    }
    c->pending.clear();
    c->evused = 0;
}

static int markers_inside(fribo_ctx* c)
{
    // lgrdistribution.F90:27-36
    const int nlm = c->F.nlm;
    if (!c->p.ibact || nlm <= 0) return FRIBO_OK;
    double ymax = -1e300, zmax = -1e300, ymin = 1e300, zmin = 1e300;
    for (int i = 0; i < nlm; ++i) {
        ymax = std::max(ymax, c->lmcc[i]);
        ymin = std::min(ymin, c->lmcc[i]);
        zmax = std::max(zmax, c->lmcc[nlm + i]);
        zmin = std::min(zmin, c->lmcc[nlm + i]);
    }
    if (ymax > c->p.rext2 || ymin < 0.0 || zmax > c->p.alx3 || zmin < 0.0)
        return set_err(FRIBO_ERR_MARKER_OUTSIDE, "a Lagrangian marker lies outside the domain");
    return FRIBO_OK;
}

static void enqueue_cfl(fribo_ctx* c)
{
    FluidDev& F = c->F;
    run(c, PC_CFL, [&] { launch_cfl(F, c->st); });
}

// phcalc_outflow (phcalc_outflow.F90:4-71).  Along z, the reference's order:
// hmn = dph * zmmt^T, one j-line per z-mode, dph = V * zmm^T.  Along y:
// hmn = Yinv * dph, one k-line per y-mode, dph = Y * V, whose epilogue leaves
// the partial sums of the constant that restores the z-order gauge (k_upd
// adds them up and subtracts the constant).
static void enqueue_poisson(fribo_ctx* c)
{
    FluidDev& F = c->F;
    hipStream_t st = c->st;
    if (F.taxis == 1 && F.refine >= 0) {
        c->dph_mirror = false;
        // mixed precision: x0 = P32(rhs), then F.refine sweeps of
        // x += P32(rhs - L x) with the residual in FP64; P32 is the solve with
        // FP32 transforms and the FP64 per-mode z-lines.  The iterate lives
        // in sA, the right-hand side (for a second sweep) in sB, the final
        // solution in dph.
        double* d = F.dph + fidx(F, 1, 1);
        double* xa = F.sA + fidx(F, 1, 1);
        double* rb = F.sB + fidx(F, 1, 1);
        const size_t fbytes = (size_t)F.P * F.n3m * sizeof(double);
        if (F.refine >= 2) hipMemcpyAsync(rb, d, fbytes, hipMemcpyDeviceToDevice, st);
        for (int it = 0; it <= F.refine; ++it) {
            if (it > 0)
                run(c, PC_JP, [&] { launch_pres_residual(F, F.sA, it == 1 ? F.dph : F.sB, F.dph, st); });
            run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
                launch_gemm32(F.n2m, F.n3m, F.n2m, F.ymmi, F.ldy, d, F.P, F.hmn, F.P, nullptr, F.S, st, nullptr,
                              nullptr, a, b);
            });
            run(c, PC_JP, [&] { launch_ysolve(F, st, false); });
            const bool last = it == F.refine;
            run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
                launch_gemm32(F.n2m, F.n3m, F.n2m, F.ymm, F.ldy, F.hmn, F.P, last ? d : xa, F.P,
                              it > 0 ? xa : nullptr, F.S, st, last ? F.zg : nullptr, last ? F.gpart : nullptr, a, b);
            });
        }
        return;
    }
    if (F.taxis == 1 && F.ysh > 0) {
        // mirror form: k_div left the pair butterfly of the divergence;
        // block-diagonal transforms with half the flops, the per-mode
        // z-lines in the mirror form's mode order; dph is left as the
        // backward transform's result, whose pair butterfly k_upd takes as
        // it reads it (the gauge partials come from rows 1 and n2m)
        double* d = F.dph + fidx(F, 1, 1);
        FluidDev Fm = F;
        Fm.wy = F.wys;
        Fm.ynull = F.ynulls;
        YFuse yf{F.wys, F.amphk, F.acphk, F.apphk, F.ysum, F.n2m, 0};
        if (F.n3m % F.ks[0].nch == 0) yf.csz = F.n3m / F.ks[0].nch;
        bool fused = false;
        run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
            fused = launch_gemm_nn(F.n2m, F.n3m, F.n2m, F.ysmi, F.ldy, d, F.P, F.hmn, F.P, F.S, st, nullptr, nullptr,
                                   a, b, &yf, F.ysh);
        });
        run(c, PC_JP, [&] { launch_ysolve(Fm, st, fused); });
        run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
            launch_gemm_nn(F.n2m, F.n3m, F.n2m, F.ysm, F.ldy, F.hmn, F.P, d, F.P, F.S, st, F.zg, F.gpart, a, b,
                           nullptr, F.ysh);
        });
        c->dph_mirror = true;
        return;
    }
    if (F.taxis == 1) {
        c->dph_mirror = false;
        // the chunk pass of the per-mode z-lines runs in the forward
        // transform's epilogue when the line partition has equal chunks
        YFuse yf{F.wy, F.amphk, F.acphk, F.apphk, F.ysum, F.n2m, 0};
        if (F.n3m % F.ks[0].nch == 0) yf.csz = F.n3m / F.ks[0].nch;
        bool fused = false;
        run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
            fused = launch_gemm_nn(F.n2m, F.n3m, F.n2m, F.ymmi, F.ldy, F.dph + fidx(F, 1, 1), F.P, F.hmn, F.P, F.S,
                                   st, nullptr, nullptr, a, b, &yf);
        });
        run(c, PC_JP, [&] { launch_ysolve(F, st, fused); });
        run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
            launch_gemm_nn(F.n2m, F.n3m, F.n2m, F.ymm, F.ldy, F.hmn, F.P, F.dph + fidx(F, 1, 1), F.P, F.S, st, F.zg,
                           F.gpart, a, b);
        });
        return;
    }
    c->dph_mirror = false;
    run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
        launch_gemm_nt(F.n2m, F.n3m, F.n3m, F.dph + fidx(F, 1, 1), F.P, F.zmmt, F.ldz, F.hmn, F.P, F.S, st, a, b);
    });
    run(c, PC_JP, [&] { launch_jsolve_poisson(F, st); });
    run_ev(c, PC_GEMM, [&](hipEvent_t a, hipEvent_t b) {
        launch_gemm_nt(F.n2m, F.n3m, F.n3m, F.hmn, F.P, F.zmm, F.ldz, F.dph + fidx(F, 1, 1), F.P, F.S, st, a, b);
    });
}

// Substep n.  fold: the previous substep's pressure mean runs inside this
// substep's k_bc (no kernel of its own); last: the mean runs after k_upd.
static void enqueue_substep(fribo_ctx* c, int n, bool fold, bool last)
{
    FluidDev& F = c->F;
    hipStream_t st = c->st;
    run(c, PC_BC, [&] { launch_bc(F, n, fold, st); });
    run(c, PC_RHS, [&] { launch_rhs(F, st); });
    if (F.ibact && F.nlm > 0) {
        if (n == 1) run(c, PC_MLS_SETUP, [&] { launch_mls_setup(F, st); });
        run(c, PC_MLS_FORCE, [&] { launch_mls_force(F, st); });
    }
    run(c, PC_JH, [&] { launch_jsolve_helm(F, st); });
    run(c, PC_K, [&] { launch_ksolve(F, st); });
    run(c, PC_DIV, [&] { launch_div(F, st); });
    enqueue_poisson(c);
    run(c, PC_UPD, [&] { launch_upd(F, st); });
    if (last) run(c, PC_MEAN, [&] { launch_mean(F, st); });
}

// Step inputs (dt, dt_o, ns) and a cleared error flag, one small copy.
static int write_inputs(fribo_ctx* c, int ns, double dt, double dt_o)
{
    c->in->dt = dt;
    c->in->dto = dt_o;
    c->in->ns = ns;
    c->in->err = 0;
    HIPCHK(hipMemcpyAsync(&c->F.S->in_dt, c->in, sizeof(fribo_ctx::StepIn), hipMemcpyHostToDevice, c->st));
    return FRIBO_OK;
}

// cfl: begin with k_cfl (not already enqueued behind the previous step).
static void enqueue_step(fribo_ctx* c, bool cfl)
{
    if (cfl) enqueue_cfl(c);
    for (int n = 1; n <= c->F.nsst; ++n) enqueue_substep(c, n, n > 1, n == c->F.nsst);
}

// Capture the whole step once; every kernel reads its per-step inputs from
// device memory, so the same graph replays for every step.
static int ensure_graph(fribo_ctx* c)
{
    if (c->gexec && c->graph_nlm == c->F.nlm) return FRIBO_OK;
    if (c->gexec) {
        hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    const bool prof = c->profiling;
    c->profiling = false;
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, false);
    hipError_t e = hipStreamEndCapture(c->st, &g);
    c->profiling = prof;
    if (e != hipSuccess) return set_err(FRIBO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(FRIBO_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    c->graph_nlm = c->F.nlm;
    return FRIBO_OK;
}

static int finish(fribo_ctx* c, double* dt, double* cflm, bool next_cfl = false)
{
    int e = read_scalars(c, next_cfl);
    if (e) return e;
    HIPCHK(hipGetLastError());
    if (!c->pending.empty()) collect_profile(c);
    if (dt) *dt = c->hs.dt;
    if (cflm) *cflm = c->hs.cflm;
    switch (c->hs.err) {
    case 0: return FRIBO_OK;
    case FRIBO_E_DT_TOO_SMALL: return set_err(FRIBO_ERR_DT_TOO_SMALL, "Too small time step size");
    case FRIBO_E_CFL_TOO_LARGE: return set_err(FRIBO_ERR_CFL_TOO_LARGE, "Too large peak CFL for stability");
    case FRIBO_E_PARTITION_UNITY: return set_err(FRIBO_ERR_PARTITION_UNITY, "partition of unity is not satisfied");
    default: return set_err(FRIBO_ERR_HIP, "device error flag " + std::to_string(c->hs.err));
    }
}

extern "C" int fribo_fluid_tsch(fribo_handle h, int32_t ns, double* dt, double dt_o, double* cflm)
{
    if (!h || !dt) return set_err(FRIBO_ERR_ARG, "null argument");
    int e = markers_inside(h);
    if (e) return e;
    if ((e = write_inputs(h, ns, *dt, dt_o))) return e;
    const bool cfl = !h->cfl_ready;
    h->cfl_ready = false;
    if (h->use_graph && !h->profiling) {
        if ((e = ensure_graph(h))) return e;
        if (cfl) enqueue_cfl(h);
        HIPCHK(hipGraphLaunch(h->gexec, h->st));
    } else {
        enqueue_step(h, cfl);
    }
    // with per-class profiling the next step's k_cfl stays in that step
    return finish(h, dt, cflm, !h->profiling);
}

extern "C" int fribo_set_graph(fribo_handle h, int32_t on)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    h->use_graph = on != 0;
    return FRIBO_OK;
}

extern "C" int fribo_cfl_dt(fribo_handle h, int32_t ns, double* dt, double dt_o, double* cflm)
{
    if (!h || !dt) return set_err(FRIBO_ERR_ARG, "null argument");
    int e = write_inputs(h, ns, *dt, dt_o);
    if (e) return e;
    enqueue_cfl(h);
    // finalize the CFL / dt control without running a substep: launch the
    // boundary kernel for substep 1 (it also prepares substep 1's factors)
    run(h, PC_BC, [&] { launch_bc(h->F, 1, false, h->st); });
    return finish(h, dt, cflm);
}

// Substep n after fribo_cfl_dt.  folded: the launch sequence of
// fribo_fluid_tsch (substep n's k_bc also runs substep n-1's pressure mean and
// boundary update; only the last substep ends with the mean kernel);
// otherwise every substep ends with the mean kernel.
static int substep(fribo_handle h, int n, bool folded)
{
    if (!h || n < 1 || n > h->F.nsst) return set_err(FRIBO_ERR_ARG, "bad substep");
    int e = 0;
    if (n == 1 && (e = markers_inside(h))) return e;
    FluidDev& F = h->F;
    hipStream_t st = h->st;
    const bool mean = !folded || n == F.nsst;
    // substep 1's boundary kernel already ran in fribo_cfl_dt; for n == 1 run
    // the body with the stored dt (launch_bc with n > 1 semantics would
    // redo boucdq), so the body starts after k_bc.
    if (n == 1) {
        run(h, PC_RHS, [&] { launch_rhs(F, st); });
        if (F.ibact && F.nlm > 0) {
            run(h, PC_MLS_SETUP, [&] { launch_mls_setup(F, st); });
            run(h, PC_MLS_FORCE, [&] { launch_mls_force(F, st); });
        }
        run(h, PC_JH, [&] { launch_jsolve_helm(F, st); });
        run(h, PC_K, [&] { launch_ksolve(F, st); });
        run(h, PC_DIV, [&] { launch_div(F, st); });
        enqueue_poisson(h);
        run(h, PC_UPD, [&] { launch_upd(F, st); });
        if (mean) run(h, PC_MEAN, [&] { launch_mean(F, st); });
    } else {
        enqueue_substep(h, n, folded, mean);
    }
    return finish(h, nullptr, nullptr);
}

extern "C" int fribo_substep(fribo_handle h, int32_t n, int32_t ns)
{
    (void)ns;
    return substep(h, n, false);
}

extern "C" int fribo_substep_folded(fribo_handle h, int32_t n, int32_t ns)
{
    (void)ns;
    return substep(h, n, true);
}

// ---------------------------------------------------------------------------
// Named array access
extern "C" long fribo_array_size(fribo_handle h, const char* name)
{
    if (!h || !name) return 0;
    const FluidDev& F = h->F;
    std::string s(name);
    const long g = (long)F.n2 * (F.n3m + 6), r = (long)F.n2 * F.n3m;
    if (s == "q2" || s == "q3" || s == "pr" || s == "dph" || s == "sA" || s == "sB") return g;
    if (s == "hmn") return (long)F.n2m * F.n3m;
    if (s == "ru2" || s == "ru3" || s == "rhs2" || s == "rhs3") return r;
    if (s == "fey" || s == "fez") return (long)F.n2 * (F.n3m + 5);
    if (s == "qb2s" || s == "qb3s" || s == "qb2n" || s == "qb3n" || s == "dqb2s" || s == "dqb3s" || s == "dqb2n" ||
        s == "dqb3n" || s == "dq2x2o" || s == "dq3x2o")
        return F.n2;
    for (int a = 0; a < A_COUNT; ++a)
        if (s == kArrNames[a]) return is_zdir(a) ? (a == A_WR ? F.n3m : F.n3) : F.n2;
    if (s == "lmind") return 4L * F.nlm;
    if (s == "eulins") return 4L * FRIBO_DIMSD * F.nlm;
    if (s == "phi" || s == "phipred") return 2L * FRIBO_DIMSD * F.nlm;
    if (s == "cl") return 2L * F.nlm;
    if (s == "scalars") return 12;
    return 0;
}

extern "C" int fribo_get_array(fribo_handle h, const char* name, void* out, long n)
{
    if (!h || !name || !out) return set_err(FRIBO_ERR_ARG, "null argument");
    long need = fribo_array_size(h, name);
    if (need == 0 || n < need) return set_err(FRIBO_ERR_ARG, std::string("unknown array or short buffer: ") + name);
    FluidDev& F = h->F;
    std::string s(name);
    int e = 0;
    HIPCHK(hipStreamSynchronize(h->st));
    if (s == "scalars") {
        if ((e = read_scalars(h))) return e;
        double* o = (double*)out;
        const DevScalars& S = h->hs;
        double v[12] = {S.dt, S.beta, S.al, S.ga, S.ro, S.cflm, S.q3avg, S.qinf, S.qout, F.dx2, F.dx3, F.ren};
        memcpy(o, v, sizeof(v));
        return FRIBO_OK;
    }
    if (s == "q2" || s == "q3" || s == "dph" || s == "sA" || s == "sB") {
        double* src = s == "q2" ? F.q2 : (s == "q3" ? F.q3 : (s == "dph" ? F.dph : (s == "sA" ? F.sA : F.sB)));
        e = get_field(h, (double*)out, src, -2, F.nk);
        if (!e && s == "dph" && h->dph_mirror) {
            // the pair butterfly of the mirror form, which k_upd applies as it reads dph
            double* o = (double*)out;
            for (int k = 1; k <= F.n3m; ++k) {
                double* col = o + (size_t)(k + 2) * F.n2;
                for (int j = 0; j < F.n2m / 2; ++j) {
                    const double a = col[j], b = col[F.n2m - 1 - j];
                    col[j] = a + b;
                    col[F.n2m - 1 - j] = a - b;
                }
            }
        }
        if (!e && s == "dph" && F.taxis == 1) {
            if ((e = read_scalars(h))) return e;
            const double off = h->hs.dphoff;
            double* o = (double*)out;
            for (int k = 1; k <= F.n3m; ++k)
                for (int j = 1; j <= F.n2m; ++j) {
                    double& v = o[(size_t)(j - 1) + (size_t)(k + 2) * F.n2];
                    v = v - off;
                }
        }
    } else if (s == "hmn") {
        HIPCHK(hipMemcpy2DAsync(out, F.n2m * sizeof(double), F.hmn, F.P * sizeof(double), F.n2m * sizeof(double),
                                F.n3m, hipMemcpyDeviceToHost, h->st));
    } else if (s == "pr") {
        return fribo_get_state(h, nullptr, nullptr, (double*)out, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr);
    } else if (s == "ru2" || s == "ru3" || s == "rhs2" || s == "rhs3") {
        double* src = s == "ru2" ? F.ru2 : (s == "ru3" ? F.ru3 : (s == "rhs2" ? F.rhs2 : F.rhs3));
        e = get_field(h, (double*)out, src, 1, F.n3m);
    } else if (s == "fey" || s == "fez") {
        const int d = (s == "fey") ? 0 : 1;
        double* o = (double*)out;
        std::fill(o, o + need, 0.0);
        const int nreal = F.nlm * FRIBO_DIMSD;
        if (F.nlm > 0 && F.ibact) {
            // segment heads of the sorted (cell, entry) keys carry the cell value
            std::vector<unsigned long long> key(nreal);
            std::vector<double> val(nreal);
            HIPCHK(hipMemcpy(key.data(), F.keys + (size_t)d * F.maxkeys, nreal * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(val.data(), F.segval + (size_t)d * F.maxkeys, nreal * sizeof(double),
                             hipMemcpyDeviceToHost));
            for (int i = 0; i < nreal; ++i) {
                if (key[i] == ~0ull) break;
                const unsigned long long cell = key[i] >> 20;
                if (i > 0 && (key[i - 1] >> 20) == cell) continue;
                const int j = (int)(cell % F.n2) + 1, k = (int)(cell / F.n2) + 1;
                o[(size_t)(j - 1) + (size_t)(k + 2) * F.n2] = val[i];
            }
        }
        return FRIBO_OK;
    } else if (s == "lmind" || s == "eulins") {
        int* src = s == "lmind" ? F.lmind : F.eulins;
        HIPCHK(hipMemcpy(out, src, need * sizeof(int), hipMemcpyDeviceToHost));
    } else if (s == "phi" || s == "phipred" || s == "cl") {
        double* src = s == "phi" ? F.phi : (s == "phipred" ? F.phipred : F.cl);
        HIPCHK(hipMemcpy(out, src, need * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        double* bsrc = nullptr;
        const char* bn[] = {"qb2s", "qb3s", "qb2n", "qb3n", "dqb2s", "dqb3s", "dqb2n", "dqb3n", "dq2x2o", "dq3x2o"};
        double* bd[] = {F.qb2s, F.qb3s, F.qb2n, F.qb3n, F.dqb2s, F.dqb3s, F.dqb2n, F.dqb3n, F.dq2x2o, F.dq3x2o};
        for (int i = 0; i < 10; ++i)
            if (s == bn[i]) bsrc = bd[i];
        if (bsrc) {
            HIPCHK(hipMemcpy(out, bsrc + 1, F.n2 * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            for (int a = 0; a < A_COUNT; ++a)
                if (s == kArrNames[a]) memcpy(out, h->h1[a].data() + 1, need * sizeof(double));
        }
    }
    if (e) return e;
    HIPCHK(hipStreamSynchronize(h->st));
    return FRIBO_OK;
}

extern "C" int fribo_sync(fribo_handle h)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->st));
    return FRIBO_OK;
}

extern "C" const char* fribo_last_error(void) { return g_err.c_str(); }

static int test_gemm(bool nn, int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                     int32_t ldb, double* C, int32_t ldc)
{
    if (M <= 0 || N <= 0 || K <= 0 || lda < M || ldb < (nn ? K : N) || ldc < M || !A || !B || !C)
        return set_err(FRIBO_ERR_ARG, "bad gemm arguments");
    const size_t nb = (size_t)ldb * (nn ? N : K);
    double *dA, *dB, *dC;
    HIPCHK(hipMalloc(&dA, (size_t)lda * K * sizeof(double)));
    HIPCHK(hipMalloc(&dB, nb * sizeof(double)));
    HIPCHK(hipMalloc(&dC, (size_t)ldc * N * sizeof(double)));
    HIPCHK(hipMemcpy(dA, A, (size_t)lda * K * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dB, B, nb * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dC, C, (size_t)ldc * N * sizeof(double), hipMemcpyHostToDevice));
    if (nn) launch_gemm_nn(M, N, K, dA, lda, dB, ldb, dC, ldc, nullptr, 0);
    else launch_gemm_nt(M, N, K, dA, lda, dB, ldb, dC, ldc, nullptr, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(C, dC, (size_t)ldc * N * sizeof(double), hipMemcpyDeviceToHost));
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    return FRIBO_OK;
}

extern "C" int fribo_test_gemm(int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                               int32_t ldb, double* C, int32_t ldc)
{
    return test_gemm(false, M, N, K, A, lda, B, ldb, C, ldc);
}

extern "C" int fribo_test_gemm_nn(int32_t M, int32_t N, int32_t K, const double* A, int32_t lda, const double* B,
                                  int32_t ldb, double* C, int32_t ldc)
{
    return test_gemm(true, M, N, K, A, lda, B, ldb, C, ldc);
}

// Streaming copy of a known byte count: the yardstick that turns the HBM
// counters (FETCH_SIZE / WRITE_SIZE) of the step's kernels into bytes, and
// the achievable-bandwidth reference.  One double per lane, 8192 workgroups
// of 256 over a grid-stride loop: the fastest of the simple copies measured
// on MI355X (5.9-6.05 TB/s on 256 MiB; two doubles per lane, unrolled or
// non-temporal forms ran at 4.6-5.6).
__global__ void __launch_bounds__(256) k_plain_copy(const double* __restrict__ a, double* __restrict__ b, size_t n)
{
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) b[i] = a[i];
}

extern "C" int fribo_test_copy(int64_t bytes, int32_t reps, double* ms_per_copy)
{
    if (bytes <= 0 || bytes % 8 != 0 || reps <= 0) return set_err(FRIBO_ERR_ARG, "bad copy arguments");
    double *a, *b;
    HIPCHK(hipMalloc(&a, (size_t)bytes));
    HIPCHK(hipMalloc(&b, (size_t)bytes));
    HIPCHK(hipMemset(a, 0, (size_t)bytes));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const size_t n = (size_t)bytes / 8;
    hipLaunchKernelGGL(k_plain_copy, dim3(8192), dim3(256), 0, 0, a, b, n);   // warm-up
    HIPCHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_plain_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_copy) *ms_per_copy = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return FRIBO_OK;
}

// loads_transfer_c (loads_transfer.F90:4-162) on the state the last step
// left: probes at hpr = ddmin*kprobe along lmnor, where cordin already scaled
// ddmin by kprobe (cordin.F90:231-233, loads_transfer.F90:41).
extern "C" int fribo_loads_transfer(fribo_handle h, double kprobe, const double* lmnor, const double* lmacc,
                                    double* press, double* tau)
{
    if (!h || !lmnor || !lmacc || !press || !tau) return set_err(FRIBO_ERR_ARG, "null argument");
    FluidDev& F = h->F;
    if (F.nlm <= 0) return set_err(FRIBO_ERR_ARG, "no markers set");
    const size_t nb = (size_t)F.nlm * sizeof(double);
    HIPCHK(hipMemcpyAsync(F.lmnor, lmnor, 2 * nb, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(F.lmacc, lmacc, 2 * nb, hipMemcpyHostToDevice, h->st));
    const double hpr = (h->dmin * kprobe) * kprobe;
    launch_loads(F, hpr, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(press, F.press, nb, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(tau, F.tau, 2 * nb, hipMemcpyDeviceToHost, h->st));
    int e = read_scalars(h);
    if (e) return e;
    if (h->hs.err) return set_err(h->hs.err, "MLS partition of unity violated at a load probe");
    return FRIBO_OK;
}

// fribo_loads_transfer with every device buffer of the call (probe inputs,
// loads, per-marker work records) in the static scratch of the handle's
// device: no allocation.  Same kernels, same results.
extern "C" int fribo_loads_transfer_static(fribo_handle h, double kprobe, const double* lmnor, const double* lmacc,
                                           double* press, double* tau)
{
    if (!h || !lmnor || !lmacc || !press || !tau) return set_err(FRIBO_ERR_ARG, "null argument");
    if (h->F.nlm <= 0) return set_err(FRIBO_ERR_ARG, "no markers set");
    if (h->F.nlm > FRIBO_STATIC_LOADS_MAX_NLM)
        return set_err(FRIBO_ERR_LIMIT, "fribo_loads_transfer_static: " + std::to_string(h->F.nlm) +
                                            " markers, the static scratch holds at most " +
                                            std::to_string(FRIBO_STATIC_LOADS_MAX_NLM));
    if (!h->scratch) return set_err(FRIBO_ERR_HIP, "static scratch address not set");
    FluidDev G = h->F;
    const size_t n = (size_t)G.nlm;
    G.lmnor = h->scratch;
    G.lmacc = G.lmnor + 2 * n;
    G.press = G.lmacc + 2 * n;
    G.tau = G.press + n;
    G.lwork = G.tau + 2 * n;
    const size_t nb = n * sizeof(double);
    HIPCHK(hipMemcpyAsync(G.lmnor, lmnor, 2 * nb, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(G.lmacc, lmacc, 2 * nb, hipMemcpyHostToDevice, h->st));
    const double hpr = (h->dmin * kprobe) * kprobe;
    launch_loads(G, hpr, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(press, G.press, nb, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(tau, G.tau, 2 * nb, hipMemcpyDeviceToHost, h->st));
    int e = read_scalars(h);
    if (e) return e;
    if (h->hs.err) return set_err(h->hs.err, "MLS partition of unity violated at a load probe");
    return FRIBO_OK;
}

// checkbcs (ckbcs.F90:5-86) on the state the last step left.
extern "C" int fribo_checkbcs(fribo_handle h, double* bcerrv, double* bcerrw, double* nbcv, double* nbcw)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    FluidDev& F = h->F;
    if (F.nlm <= 0) return set_err(FRIBO_ERR_ARG, "no markers set");
    launch_checkbcs(F, h->nb, h->st);
    HIPCHK(hipGetLastError());
    const size_t nb = (size_t)F.nlm * sizeof(double);
    if (bcerrv) HIPCHK(hipMemcpyAsync(bcerrv, F.bcerr, nb, hipMemcpyDeviceToHost, h->st));
    if (bcerrw) HIPCHK(hipMemcpyAsync(bcerrw, F.bcerr + F.nlm, nb, hipMemcpyDeviceToHost, h->st));
    double v[2];
    HIPCHK(hipMemcpyAsync(v, h->nb, sizeof(v), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (nbcv) *nbcv = v[0];
    if (nbcw) *nbcw = v[1];
    return FRIBO_OK;
}

// ttime's diagnostics (ttime.F90:29-69,136) of the current state.
extern "C" int fribo_diagnostics(fribo_handle h, double* out)
{
    if (!h || !out) return set_err(FRIBO_ERR_ARG, "null argument");
    double* res = h->dpart + 8 * 256;
    launch_diag(h->F, h->dpart, h->vort_k, res, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, res, 9 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return FRIBO_OK;
}

// Rollback copy for iterated (strong) coupling: q2, q3, pr, ru2, ru3 (the
// whole device arrays) -- the q2_o, q3_o, pr_o, ru2_o, ru3_o that the
// reference declares "for strong coupling" and never uses
// (declarations.F90:155-159) -- plus the ten outflow arrays and the step
// scalars a repeated step also needs.  Device to device, on the library's
// stream; allocated on first use.
static size_t snap_doubles(const FluidDev& F)
{
    return 5 * (size_t)F.P * F.nk + 10 * (size_t)(F.n2 + 2) + (sizeof(DevScalars) + 7) / 8;
}

static int snap_copy(fribo_ctx* h, bool save)
{
    FluidDev& F = h->F;
    const size_t fsz = (size_t)F.P * F.nk * sizeof(double), bsz = (size_t)(F.n2 + 2) * sizeof(double);
    double* fl[5] = {F.q2, F.q3, F.pr, F.ru2, F.ru3};
    double* bc[10] = {F.qb2s, F.qb3s, F.qb2n, F.qb3n, F.dqb2s, F.dqb3s, F.dqb2n, F.dqb3n, F.dq2x2o, F.dq3x2o};
    char* p = (char*)h->snap;
    for (int i = 0; i < 5; ++i, p += fsz)
        HIPCHK(hipMemcpyAsync(save ? (void*)p : (void*)fl[i], save ? (void*)fl[i] : (void*)p, fsz,
                              hipMemcpyDeviceToDevice, h->st));
    for (int i = 0; i < 10; ++i, p += bsz)
        HIPCHK(hipMemcpyAsync(save ? (void*)p : (void*)bc[i], save ? (void*)bc[i] : (void*)p, bsz,
                              hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(save ? (void*)p : (void*)F.S, save ? (void*)F.S : (void*)p, sizeof(DevScalars),
                          hipMemcpyDeviceToDevice, h->st));
    return FRIBO_OK;
}

extern "C" int fribo_save_state(fribo_handle h)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    if (!h->snap) {
        int e = dalloc(h, &h->snap, snap_doubles(h->F));
        if (e) return e;
    }
    int e = snap_copy(h, true);
    if (e) return e;
    h->snap_valid = true;
    return FRIBO_OK;
}

extern "C" int fribo_restore_state(fribo_handle h)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    if (!h->snap_valid) return set_err(FRIBO_ERR_ARG, "no saved state");
    h->cfl_ready = false;
    return snap_copy(h, false);
}

extern "C" int fribo_set_transform(fribo_handle h, int32_t axis)
{
    if (!h || axis < -1 || axis > 1) return set_err(FRIBO_ERR_ARG, "bad transform axis");
    if (axis == 1 && !h->yavail)
        return set_err(FRIBO_ERR_LIMIT, "no transform along y for this grid (needs n2-1 < n3-1 and Neumann null modes)");
    const int want = axis == -1 ? (h->yavail ? 1 : 0) : axis;
    if (want != h->F.taxis && h->batch) return set_err(FRIBO_ERR_ARG, "instance belongs to a batch");
    if (want != h->F.taxis) {
        HIPCHK(hipStreamSynchronize(h->st));
        h->F.taxis = want;
        if (want != 1) h->F.refine = -1;
        apply_ysh(h);
        const double zero = 0.0;
        HIPCHK(hipMemcpy(&h->F.S->dphoff, &zero, sizeof(double), hipMemcpyHostToDevice));
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
            h->graph_nlm = -1;
        }
    }
    return FRIBO_OK;
}

// Mirror form of the transform along y (no counterpart in the reference):
// on = 1 selects it (the default when the grid is symmetric about the
// channel centre), 0 the plain dense form.  It applies to the FP64
// transforms along y (the FP32 ones keep the dense form).
extern "C" int fribo_set_ymirror(fribo_handle h, int32_t on)
{
    if (!h || on < 0 || on > 1) return set_err(FRIBO_ERR_ARG, "on must be 0 or 1");
    if (on && !h->ysh_avail)
        return set_err(FRIBO_ERR_LIMIT, "no mirror form: the y grid is not symmetric about the centre, or n2-1 is odd");
    if (h->batch) return set_err(FRIBO_ERR_ARG, "instance belongs to a batch");
    if ((on != 0) != h->ymirror_on) {
        HIPCHK(hipStreamSynchronize(h->st));
        h->ymirror_on = on != 0;
        apply_ysh(h);
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
            h->graph_nlm = -1;
        }
    }
    return FRIBO_OK;
}

extern "C" int fribo_get_ymirror(fribo_handle h)
{
    return h ? (h->ymirror_on && h->ysh_avail > 0 ? 1 : 0) : -1;
}

extern "C" int fribo_get_transform(fribo_handle h)
{
    return h ? h->F.taxis : -1;
}

extern "C" int fribo_set_ksolve(fribo_handle h, int32_t form)
{
    if (!h || form < 0 || form > 1) return set_err(FRIBO_ERR_ARG, "bad k-line solve form");
    if (h->batch) return set_err(FRIBO_ERR_ARG, "instance belongs to a batch");
    if (form == 1 && !ksolve_fused_fits(h->F))
        return set_err(FRIBO_ERR_LIMIT, "k-lines too long for the one-pass solve (chunks of more than 32 rows)");
    if (form != h->F.kform) {
        HIPCHK(hipStreamSynchronize(h->st));
        h->F.kform = form;
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
            h->graph_nlm = -1;
        }
    }
    return FRIBO_OK;
}

extern "C" int fribo_get_ksolve(fribo_handle h)
{
    return h ? h->F.kform : -1;
}

extern "C" int fribo_set_profiling(fribo_handle h, int32_t on)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    h->profiling = on != 0;
    h->prof_mask = (on == 2) ? (1u << PC_GEMM) : ~0u;
    return FRIBO_OK;
}

extern "C" int fribo_get_profile(fribo_handle h, double* ms, int32_t* launches, int32_t nclass)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    for (int i = 0; i < nclass && i < FRIBO_NPROF; ++i) {
        if (ms) ms[i] = h->prof_ms[i];
        if (launches) launches[i] = h->prof_n[i];
    }
    return FRIBO_OK;
}

extern "C" int fribo_reset_profile(fribo_handle h)
{
    if (!h) return set_err(FRIBO_ERR_ARG, "null handle");
    for (int i = 0; i < FRIBO_NPROF; ++i) {
        h->prof_ms[i] = 0.0;
        h->prof_n[i] = 0;
    }
    return FRIBO_OK;
}

extern "C" const char* fribo_profile_name(int32_t cls)
{
    return (cls >= 0 && cls < FRIBO_NPROF) ? kProfNames[cls] : "";
}

// ---------------------------------------------------------------------------
// Batches of independent instances (no counterpart in the reference: many
// runs of the fluid step, on uneven grids, advanced together).  Every stage
// of the step is one launch over all instances (fribo_batch.h); the
// instances keep their own device state, created by fribo_init and filled
// by the single-instance calls, and share the batch's stream.

struct BatchOut {
    double dt, cflm;
    int err, pad;
};

__global__ __launch_bounds__(64) void k_batch_inputs(const FluidDev* __restrict__ F,
                                                      const double* __restrict__ in, int n)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    DevScalars* S = F[i].S;
    S->in_dt = in[3 * i];
    S->in_dto = in[3 * i + 1];
    const long long w = __double_as_longlong(in[3 * i + 2]);
    S->in_ns = (int)(w & 0xffffffff);
    S->err = (int)(w >> 32);
}

__global__ __launch_bounds__(64) void k_batch_outputs(const FluidDev* __restrict__ F, BatchOut* __restrict__ out,
                                                       int n)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const DevScalars* S = F[i].S;
    out[i] = BatchOut{S->dt, S->cflm, S->err, 0};
}

// marker kinematics of all instances from one staged copy: instance
// blockIdx.y takes n5[i] doubles from src + off[i] into its lmcc buffer
__global__ __launch_bounds__(256) void k_batch_markers(const FluidDev* __restrict__ F, const long* __restrict__ off,
                                                       const double* __restrict__ src)
{
    const int i = blockIdx.y;
    const long n5 = 5 * (long)F[i].nlm;
    double* dst = F[i].lmcc;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n5; q += (long)gridDim.x * 256) dst[q] = src[off[i] + q];
}

struct BTable {
    std::vector<int> dims;
    std::vector<GemmProb> probs;
    int* dbuf = nullptr;         // off [n+1], dims [3n]
    GemmProb* dprob = nullptr;
};

struct fribo_batch_ctx {
    std::vector<fribo_handle> inst;
    std::vector<int> gen;              // marker generation of each instance at the last refresh
    BatchSet B;
    hipStream_t st = nullptr;
    BTable tab[BK_COUNT];
    int jclass = 8;                    // j-line chunk class of the batch
    int gemm_fixed = -1;               // tile shape set by the caller or the autotuner (-1: default)
    double* in_h = nullptr;            // pinned step inputs [3n]
    double* in_d = nullptr;
    BatchOut* out_h = nullptr;         // pinned step outputs [n]
    BatchOut* out_d = nullptr;
    // batched marker upload
    double* mk_h = nullptr;
    double* mk_d = nullptr;
    long* mk_off = nullptr;
    long mk_cap = 0;
    hipEvent_t ev_mk = nullptr;
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    bool warm = false;                 // an eager step ran with the current tables
};

static void batch_drop_graph(fribo_batch_ctx* b);

static BGrid<FluidDev> batch_grid(BatchSet* B, int key, const std::vector<int>& dims)
{
    fribo_batch_ctx* b = (fribo_batch_ctx*)B->owner;
    BTable& T = b->tab[key];
    if (!T.dbuf || T.dims != dims) {
        if (T.dbuf) hipFree(T.dbuf);
        T.dbuf = nullptr;
        std::vector<int> h(B->n + 1 + dims.size());
        int acc = 0;
        for (int i = 0; i < B->n; ++i) {
            h[i] = acc;
            acc += dims[3 * i] * dims[3 * i + 1] * dims[3 * i + 2];
        }
        h[B->n] = acc;
        std::copy(dims.begin(), dims.end(), h.begin() + B->n + 1);
        if (hipMalloc(&T.dbuf, h.size() * sizeof(int)) != hipSuccess ||
            hipMemcpy(T.dbuf, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            T.dbuf = nullptr;
            return BGrid<FluidDev>{B->dF, nullptr, nullptr, 0};
        }
        T.dims = dims;
        batch_drop_graph(b);
    }
    return BGrid<FluidDev>{B->dF, T.dbuf, T.dbuf + B->n + 1, B->n};
}

static bool same_prob(const GemmProb& a, const GemmProb& b)
{
    return a.M == b.M && a.N == b.N && a.K == b.K && a.lda == b.lda && a.ldb == b.ldb && a.ldc == b.ldc &&
           a.A == b.A && a.B == b.B && a.C == b.C && a.S == b.S && a.zg == b.zg && a.gpart == b.gpart &&
           a.yf.wy == b.yf.wy && a.yf.amphk == b.yf.amphk && a.yf.acphk == b.yf.acphk && a.yf.apphk == b.yf.apphk &&
           a.yf.ysum == b.yf.ysum && a.yf.n2m == b.yf.n2m && a.yf.csz == b.yf.csz && a.ks == b.ks;
}

static bool same_probs(const std::vector<GemmProb>& a, const std::vector<GemmProb>& b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (!same_prob(a[i], b[i])) return false;
    return true;
}

// A rebuilt table invalidates the captured step, whose kernel arguments
// point to the old one.
static void batch_drop_graph(fribo_batch_ctx* b)
{
    if (b->gexec) {
        hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    b->warm = false;
}

static BGrid<GemmProb> batch_gemm(BatchSet* B, int key, const std::vector<GemmProb>& probs,
                                  const std::vector<int>& dims)
{
    fribo_batch_ctx* b = (fribo_batch_ctx*)B->owner;
    BTable& T = b->tab[key];
    const BGrid<FluidDev> g = batch_grid(B, key, dims);
    if (!g.off) return BGrid<GemmProb>{nullptr, nullptr, nullptr, 0};
    if (!T.dprob || !same_probs(T.probs, probs)) {
        if (T.dprob) hipFree(T.dprob);
        T.dprob = nullptr;
        if (hipMalloc(&T.dprob, probs.size() * sizeof(GemmProb)) != hipSuccess ||
            hipMemcpy(T.dprob, probs.data(), probs.size() * sizeof(GemmProb), hipMemcpyHostToDevice) != hipSuccess) {
            T.dprob = nullptr;
            return BGrid<GemmProb>{nullptr, nullptr, nullptr, 0};
        }
        T.probs = probs;
        batch_drop_graph(b);
    }
    return BGrid<GemmProb>{T.dprob, g.off, g.dims, B->n};
}

// Default tile shape of the batch's transforms: the one the largest
// instance would take on its own, among those every instance can run.
static int batch_default_gemm(const fribo_batch_ctx* b)
{
    int mask = 31;
    long best = -1;
    const FluidDev* big = nullptr;
    for (auto h : b->inst) {
        mask &= gemm_cfg_mask(h->F);
        const long cells = (long)h->F.n2m * h->F.n3m;
        if (cells > best) {
            best = cells;
            big = &h->F;
        }
    }
    const int M = big->n2m, N = big->n3m;
    const auto cdiv = [](int a, int c) { return (a + c - 1) / c; };
    int cfg = 0;
    if ((long)cdiv(M, 96) * cdiv(N, 128) >= 192) cfg = (mask & 8) ? 3 : 2;
    else if ((long)cdiv(M, 64) * cdiv(N, 64) >= 96) cfg = 1;
    while (cfg > 0 && !((mask >> cfg) & 1)) --cfg;
    return cfg;
}

// The instances' FluidDev as the batched kernels see them: the j-line
// staging stride of the batch's chunk class, the gauge partial count of its
// tile shape.  Uploaded as one array.
static int batch_refresh(fribo_batch_ctx* b)
{
    BatchSet& B = b->B;
    B.gemm_cfg = b->gemm_fixed >= 0 ? b->gemm_fixed : batch_default_gemm(b);
    bool fuse = B.gemm_cfg == 3 || B.gemm_cfg == 4;
    for (auto h : b->inst) fuse = fuse && gemm_fuse_ok(h->F, B.gemm_cfg);
    B.ysolve_fused = fuse;
    for (int i = 0; i < B.n; ++i) {
        FluidDev F = b->inst[i]->F;
        F.jstride = jsolve_stride_for(F, b->jclass);
        if (F.taxis != 1) F.ysh = 0;
        if (F.ysh > 0) {
            // the mirror form's mode order for the per-mode z-lines
            F.wy = F.wys;
            F.ynull = F.ynulls;
        }
        if (F.taxis == 1) F.ngpart = gemm_cfg_parts(F, B.gemm_cfg) * (F.ysh > 0 ? 2 : 1);
        B.hF[i] = F;
        b->gen[i] = b->inst[i]->mk_gen;
    }
    HIPCHK(hipMemcpyAsync(B.dF, B.hF.data(), B.n * sizeof(FluidDev), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    if (b->gexec) {
        hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    b->warm = false;
    return FRIBO_OK;
}

// The marker-staging event of an instance whose stream is about to be
// destroyed: its last record would be on that stream, so a fresh
// (unrecorded) event takes its place.
static void rearm_marker_event(fribo_ctx* h)
{
    if (!h->ev_mk) return;
    hipEventSynchronize(h->ev_mk);
    hipEventDestroy(h->ev_mk);
    h->ev_mk = nullptr;
    hipEventCreateWithFlags(&h->ev_mk, hipEventDisableTiming);
}

static void batch_free(fribo_batch_ctx* b)
{
    if (b->st) hipStreamSynchronize(b->st);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    for (auto& T : b->tab) {
        if (T.dbuf) hipFree(T.dbuf);
        if (T.dprob) hipFree(T.dprob);
    }
    if (b->B.dF) hipFree(b->B.dF);
    if (b->in_h) hipHostFree(b->in_h);
    if (b->in_d) hipFree(b->in_d);
    if (b->out_h) hipHostFree(b->out_h);
    if (b->out_d) hipFree(b->out_d);
    if (b->mk_h) hipHostFree(b->mk_h);
    if (b->mk_d) hipFree(b->mk_d);
    if (b->mk_off) hipFree(b->mk_off);
    if (b->ev_mk) hipEventDestroy(b->ev_mk);
    // the instances get streams of their own again
    for (auto h : b->inst) {
        if (h->batch != b) continue;
        rearm_marker_event(h);
        hipStreamCreate(&h->st);
        h->batch = nullptr;
    }
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

extern "C" int fribo_batch_create(const fribo_handle* hs, int32_t n, fribo_batch* out)
{
    if (!hs || !out || n < 1) return set_err(FRIBO_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    bool all_y = true;
    for (int i = 0; i < n; ++i) {
        fribo_handle h = hs[i];
        if (!h) return set_err(FRIBO_ERR_ARG, "null instance handle");
        if (h->batch) return set_err(FRIBO_ERR_ARG, "instance already belongs to a batch");
        for (int j = 0; j < i; ++j)
            if (hs[j] == h) return set_err(FRIBO_ERR_ARG, "instance listed twice");
        if (h->dev != hs[0]->dev) return set_err(FRIBO_ERR_ARG, "instances on different devices");
        if (h->F.nsst != hs[0]->F.nsst) return set_err(FRIBO_ERR_ARG, "instances with different nsst");
        if (!mls_batchable(h->F)) return set_err(FRIBO_ERR_LIMIT, "marker count beyond the batched marker order");
        all_y = all_y && h->yavail;
    }
    fribo_batch_ctx* b = new fribo_batch_ctx();
    b->inst.assign(hs, hs + n);
    b->gen.assign(n, -1);
    b->B.n = n;
    b->B.hF.resize(n);
    b->B.owner = b;
    b->B.grid_fn = batch_grid;
    b->B.gemm_fn = batch_gemm;
    for (auto h : b->inst) b->jclass = std::max(b->jclass, jsolve_class_h(h->F));
    int e = 0;
    hipError_t he = hipStreamCreate(&b->st);
    if (he == hipSuccess) he = hipMalloc(&b->B.dF, n * sizeof(FluidDev));
    if (he == hipSuccess) he = hipHostMalloc((void**)&b->in_h, 3 * (size_t)n * sizeof(double), hipHostMallocDefault);
    if (he == hipSuccess) he = hipMalloc(&b->in_d, 3 * (size_t)n * sizeof(double));
    if (he == hipSuccess) he = hipHostMalloc((void**)&b->out_h, n * sizeof(BatchOut), hipHostMallocDefault);
    if (he == hipSuccess) he = hipMalloc(&b->out_d, n * sizeof(BatchOut));
    if (he == hipSuccess) he = hipEventCreateWithFlags(&b->ev_mk, hipEventDisableTiming);
    if (he != hipSuccess) {
        e = set_err(FRIBO_ERR_HIP, hipGetErrorString(he));
        b->inst.clear();
        batch_free(b);
        return e;
    }
    // one stream for the batch and its instances; every instance takes the
    // same transform axis
    for (auto h : b->inst) {
        hipStreamSynchronize(h->st);
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
            h->graph_nlm = -1;
        }
        h->cfl_ready = false;
        fribo_set_transform(h, all_y ? 1 : 0);
        rearm_marker_event(h);
        hipStreamDestroy(h->st);
        h->st = b->st;
        h->batch = b;
    }
    if ((e = batch_refresh(b))) {
        batch_free(b);
        return e;
    }
    *out = b;
    return FRIBO_OK;
}

extern "C" int fribo_batch_destroy(fribo_batch b)
{
    if (!b) return set_err(FRIBO_ERR_ARG, "null batch");
    batch_free(b);
    return FRIBO_OK;
}

// Marker kinematics of every instance in one copy: lmcc, lmvel (nlm_i, 2)
// and lmA (nlm_i) of instance i at the offsets 5*sum(nlm_<i), in the
// reference layouts, instance after instance (lmcc_i, lmvel_i, lmA_i).  The
// instances' marker counts must be the ones of their last
// fribo_set_markers.
extern "C" int fribo_batch_set_markers(fribo_batch b, const double* mk)
{
    if (!b || !mk) return set_err(FRIBO_ERR_ARG, "null argument");
    const int n = b->B.n;
    long tot = 0;
    std::vector<long> off(n);
    for (int i = 0; i < n; ++i) {
        off[i] = tot;
        tot += 5 * (long)b->inst[i]->F.nlm;
    }
    if (tot == 0) return FRIBO_OK;
    if (tot > b->mk_cap) {
        HIPCHK(hipStreamSynchronize(b->st));
        if (b->mk_h) hipHostFree(b->mk_h);
        if (b->mk_d) hipFree(b->mk_d);
        if (b->mk_off) hipFree(b->mk_off);
        b->mk_h = nullptr;
        b->mk_d = nullptr;
        b->mk_off = nullptr;
        b->mk_cap = 0;
        HIPCHK(hipHostMalloc((void**)&b->mk_h, tot * sizeof(double), hipHostMallocDefault));
        HIPCHK(hipMalloc(&b->mk_d, tot * sizeof(double)));
        HIPCHK(hipMalloc(&b->mk_off, n * sizeof(long)));
        b->mk_cap = tot;
    }
    HIPCHK(hipEventSynchronize(b->ev_mk));
    memcpy(b->mk_h, mk, tot * sizeof(double));
    for (int i = 0; i < n; ++i) {
        fribo_ctx* h = b->inst[i];
        const long nl = h->F.nlm;
        h->lmcc.assign(mk + off[i], mk + off[i] + 2 * nl);
    }
    HIPCHK(hipMemcpyAsync(b->mk_off, off.data(), n * sizeof(long), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->mk_d, b->mk_h, tot * sizeof(double), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipEventRecord(b->ev_mk, b->st));
    for (int i = 0; i < n; ++i)
        if (b->gen[i] != b->inst[i]->mk_gen) {
            int e = batch_refresh(b);
            if (e) return e;
            break;
        }
    int mx = 1;
    for (auto h : b->inst) mx = std::max<long>(mx, (5 * (long)h->F.nlm + 255) / 256);
    hipLaunchKernelGGL(k_batch_markers, dim3(std::min(mx, 64), n), dim3(256), 0, b->st, (const FluidDev*)b->B.dF,
                       (const long*)b->mk_off, (const double*)b->mk_d);
    HIPCHK(hipGetLastError());
    return FRIBO_OK;
}

static void batch_enqueue(fribo_batch_ctx* b)
{
    BatchSet& B = b->B;
    hipStream_t st = b->st;
    const int nsst = B.hF[0].nsst;
    bool any_mk = false;
    for (const auto& F : B.hF) any_mk = any_mk || (F.ibact && F.nlm > 0);
    launch_cfl_b(B, st);
    for (int n = 1; n <= nsst; ++n) {
        launch_bc_b(B, n, n > 1, st);
        launch_rhs_b(B, st);
        if (any_mk) {
            if (n == 1) launch_mls_setup_b(B, st);
            launch_mls_force_b(B, st);
        }
        launch_jsolve_helm_b(B, st);
        launch_ksolve_b(B, st);
        launch_div_b(B, st);
        const bool fused = launch_poisson_gemm_b(B, true, st);
        if (B.hF[0].taxis == 1) launch_ysolve_b(B, st, fused);
        else launch_jsolve_poisson_b(B, st);
        launch_poisson_gemm_b(B, false, st);
        launch_upd_b(B, st);
        if (n == nsst) launch_mean_b(B, st);
    }
}

// One time step of every instance (fluid_tsch.F90:25-166 each): dt[i]
// in/out, dt_o[i] (nsst = 1), cflm[i] and err[i] out (the codes of
// fribo_fluid_tsch; an instance with an error leaves its state as its own
// step would).  Returns FRIBO_OK when the launches went through, whatever
// the instances' codes.
extern "C" int fribo_batch_step(fribo_batch b, int32_t ns, double* dt, const double* dt_o, double* cflm,
                                int32_t* err)
{
    if (!b || !dt) return set_err(FRIBO_ERR_ARG, "null argument");
    const int n = b->B.n;
    for (int i = 0; i < n; ++i)
        if (b->gen[i] != b->inst[i]->mk_gen) {
            int e = batch_refresh(b);
            if (e) return e;
            break;
        }
    for (int i = 0; i < n; ++i) {
        fribo_ctx* h = b->inst[i];
        h->cfl_ready = false;
        const int pre = markers_inside(h) ? FRIBO_E_MARKER_OUTSIDE : 0;
        b->in_h[3 * i] = dt[i];
        b->in_h[3 * i + 1] = dt_o ? dt_o[i] : dt[i];
        const long long w = ((long long)pre << 32) | (unsigned)ns;
        memcpy(&b->in_h[3 * i + 2], &w, sizeof(w));
    }
    HIPCHK(hipMemcpyAsync(b->in_d, b->in_h, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, b->st));
    hipLaunchKernelGGL(k_batch_inputs, dim3((n + 63) / 64), dim3(64), 0, b->st, (const FluidDev*)b->B.dF,
                       (const double*)b->in_d, n);
    if (b->use_graph && b->warm) {
        if (!b->gexec) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            batch_enqueue(b);
            hipError_t e = hipStreamEndCapture(b->st, &g);
            if (e != hipSuccess) return set_err(FRIBO_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return set_err(FRIBO_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
        }
        HIPCHK(hipGraphLaunch(b->gexec, b->st));
    } else {
        batch_enqueue(b);
        b->warm = true;
    }
    hipLaunchKernelGGL(k_batch_outputs, dim3((n + 63) / 64), dim3(64), 0, b->st, (const FluidDev*)b->B.dF,
                       b->out_d, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->out_h, b->out_d, n * sizeof(BatchOut), hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    for (int i = 0; i < n; ++i) {
        const long long w = *(const long long*)&b->in_h[3 * i + 2];
        const bool outside = (w >> 32) != 0;   // not stepped: dt stays as given
        if (!outside) {
            dt[i] = b->out_h[i].dt;
            b->inst[i]->dph_mirror = b->B.hF[i].ysh > 0;
        }
        if (cflm) cflm[i] = b->out_h[i].cflm;
        if (err) err[i] = b->out_h[i].err;
    }
    return FRIBO_OK;
}

extern "C" int fribo_batch_set_graph(fribo_batch b, int32_t on)
{
    if (!b) return set_err(FRIBO_ERR_ARG, "null batch");
    b->use_graph = on != 0;
    if (!b->use_graph && b->gexec) {
        hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    return FRIBO_OK;
}

// Tile shape of the batch's transforms: 0 (32x32), 1 (64x64), 2 (96x128),
// 3 (direct-to-LDS 96x128), 4 (direct-to-LDS 96x64) or -1 for the default.
extern "C" int fribo_batch_set_gemm_shape(fribo_batch b, int32_t cfg)
{
    if (!b || cfg < -1 || cfg > 4) return set_err(FRIBO_ERR_ARG, "bad tile shape");
    if (cfg >= 0)
        for (auto h : b->inst)
            if (!((gemm_cfg_mask(h->F) >> cfg) & 1))
                return set_err(FRIBO_ERR_LIMIT, "tile shape not available for every instance");
    b->gemm_fixed = cfg;
    return batch_refresh(b);
}

extern "C" int fribo_batch_get_gemm_shape(fribo_batch b)
{
    return b ? b->B.gemm_cfg : -1;
}

// Autotuner of the batch's transform tile shape: the forward transform of
// every instance (its output hmn is scratch between steps; nothing of the
// state changes) with each shape all instances can run, reps timed launches
// after one untimed, on the batch's own problems; the fastest shape is kept.
// ms[5] receives the mean time per grouped launch of shapes 0..4 (0 where a
// shape is not available).
extern "C" int fribo_batch_autotune(fribo_batch b, int32_t reps, double* ms, int32_t* best_out)
{
    if (!b || reps < 1) return set_err(FRIBO_ERR_ARG, "bad autotune arguments");
    int mask = 31;
    for (auto h : b->inst) mask &= gemm_cfg_mask(h->F);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int best = -1;
    float best_ms = 0.0f;
    const int keep = b->gemm_fixed;
    for (int cfg = 0; cfg < 5; ++cfg) {
        if (ms) ms[cfg] = ms[5 + cfg] = 0.0;
        if (!((mask >> cfg) & 1)) continue;
        b->gemm_fixed = cfg;
        int e = batch_refresh(b);
        if (e) {
            b->gemm_fixed = keep;
            return e;
        }
        // the forward transform alone, then with the chunk pass of the
        // per-mode z-lines where this shape does not run it in the
        // transform's epilogue (the time that decides)
        float tt[2] = {0.0f, 0.0f};
        for (int pass = 0; pass < 2; ++pass) {
            auto fwd = [&] {
                const bool fused = launch_poisson_gemm_b(b->B, true, b->st);
                if (pass == 1 && !fused && b->B.hF[0].taxis == 1) launch_ysolve1_b(b->B, b->st);
            };
            fwd();
            HIPCHK(hipEventRecord(e0, b->st));
            for (int r = 0; r < reps; ++r) fwd();
            HIPCHK(hipEventRecord(e1, b->st));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventElapsedTime(&tt[pass], e0, e1));
            tt[pass] /= (float)reps;
        }
        const float t = tt[1];
        if (ms) {
            ms[cfg] = tt[0];
            ms[5 + cfg] = tt[1];
        }
        if (best < 0 || t < best_ms) {
            best = cfg;
            best_ms = t;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->gemm_fixed = best;
    if (best_out) *best_out = best;
    return batch_refresh(b);
}

// Precision of the Poisson transforms (no counterpart in the reference, whose
// DGEMMs are FP64): refine < 0, FP64 (the default); refine = n >= 0, FP32
// MFMA transforms with FP32 accumulation followed by n sweeps of FP64-residual
// refinement (x += solve(rhs - L x)).  Transforms along y only.
extern "C" int fribo_set_precision(fribo_handle h, int32_t refine)
{
    if (!h || refine > 3) return set_err(FRIBO_ERR_ARG, "refine must be < 0 (FP64) or 0..3");
    if (refine >= 0 && h->F.taxis != 1)
        return set_err(FRIBO_ERR_LIMIT, "the mixed-precision solve needs the transform along y");
    if (h->batch) return set_err(FRIBO_ERR_ARG, "instance belongs to a batch");
    HIPCHK(hipStreamSynchronize(h->st));
    h->F.refine = refine < 0 ? -1 : refine;
    apply_ysh(h);
    if (h->gexec) {
        hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
        h->graph_nlm = -1;
    }
    return FRIBO_OK;
}

extern "C" int fribo_get_precision(fribo_handle h)
{
    return h ? h->F.refine : -2;
}
