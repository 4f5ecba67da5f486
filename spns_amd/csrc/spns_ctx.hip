// Host side of the MI355X HIT step: context, device memory plan, the step's
// launch sequence and the extern "C" entry points declared in include/spns.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spns.h"
#include "spns_internal.hpp"

using namespace spns;

namespace {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw Error(SPNS_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr double kPi = 3.141592653589793238462643;  // src/util.h:57

// exp(-2 pi i m/n) in long double with octant reduction (quarter and eighth
// points exact by symmetry).
void twiddle(long m, long n, double* re, double* im)
{
    m %= n;
    if (m < 0) m += n;
    long a = 8 * m;
    bool sflip = false, cflip = false, swp = false;
    if (a > 4 * n) { a = 8 * n - a; sflip = true; }
    if (a > 2 * n) { a = 4 * n - a; cflip = true; }
    if (a > n) { a = 2 * n - a; swp = true; }
    long double th = (long double)kPi * (long double)a / (4.0L * (long double)n);
    long double c = cosl(th), s = sinl(th);
    if (a == 0) { c = 1.0L; s = 0.0L; }
    if (swp) std::swap(c, s);
    if (cflip) c = -c;
    if (sflip) s = -s;
    *re = (double)c;
    *im = (double)(-s);
}

void* upload_twiddles(int L, int prec)
{
    std::vector<double> t(2 * (size_t)L);
    for (int m = 0; m < L; m++) twiddle(m, L, &t[2 * m], &t[2 * m + 1]);
    void* d = nullptr;
    if (prec == PREC_FP64) {
        HIPCHK(hipMalloc(&d, sizeof(double) * 2 * L));
        HIPCHK(hipMemcpy(d, t.data(), sizeof(double) * 2 * L, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(2 * (size_t)L);
        for (size_t i = 0; i < f.size(); i++) f[i] = (float)t[i];
        HIPCHK(hipMalloc(&d, sizeof(float) * 2 * L));
        HIPCHK(hipMemcpy(d, f.data(), sizeof(float) * 2 * L, hipMemcpyHostToDevice));
    }
    return d;
}

inline int ffreq(int a, int M) { return (a <= M / 2) ? a : a - M; }  // src/util.h:67

// Checks every address a strided launch can touch against the buffers.
void check_strided(StridedArgs& a, int A, const int64_t* in_cap, const int64_t* out_cap, int tile)
{
    static const bool plain = [] {
        const char* v = std::getenv("SPNS_STRIDED");
        return v && std::string(v) == "plain";
    }();
    a.padded_ok = tile > 0 && !plain;
    for (int j = 0; j < a.njobs; j++) {
        int64_t imax = (int64_t)(A - 1) * a.in_a_stride + (int64_t)(a.nin - 1) * a.in_row_stride + (a.B - 1);
        int64_t omax = (int64_t)(A - 1) * a.out_a_stride + (int64_t)(a.nout - 1) * a.out_row_stride + (a.B - 1);
        if (imax >= in_cap[j] || omax >= out_cap[j] || a.nin < 1 || a.nout < 1 || a.B < 1)
            throw Error(SPNS_EINVAL, "strided pass would address outside its buffers");
        if (tile > 0) {
            // the staged kernel copies whole column tiles: columns up to round_up(B, tile) - 1
            const int64_t bpad = ((int64_t)a.B + tile - 1) / tile * tile;
            if (imax - (a.B - 1) + (bpad - 1) >= in_cap[j]) a.padded_ok = 0;
        }
    }
}

}  // namespace

// ===========================================================================
// FFT_package drop-in
// ===========================================================================
struct spns_fft {
    int Mx, My, Mz, prec, device;
    void *twx = nullptr, *twy = nullptr, *twz = nullptr;
    hipStream_t st = nullptr;
    std::string err;

    void fft_x(void* c, int dir)
    {
        const int nk = Mz / 2 + 1;
        StridedArgs a{};
        a.job[0] = StridedJob{c, c, 0, 0};
        a.njobs = 1;
        a.B = nk; a.nin = Mx; a.nout = Mx;
        a.in_row_stride = nk; a.out_row_stride = nk;
        a.in_a_stride = (int64_t)Mx * nk; a.out_a_stride = (int64_t)Mx * nk;
        a.tw = twx;
        int64_t cap = (int64_t)My * Mx * nk;
        check_strided(a, My, &cap, &cap, 0);
        HIPCHK(launch_strided(Mx, prec, prec == PREC_FP64, prec == PREC_FP64, dir, a, My, st));
    }
    void fft_y(void* c, int dir)
    {
        const int nk = Mz / 2 + 1;
        StridedArgs a{};
        a.job[0] = StridedJob{c, c, 0, 0};
        a.njobs = 1;
        a.B = Mx * nk; a.nin = My; a.nout = My;
        a.in_row_stride = (int64_t)Mx * nk; a.out_row_stride = (int64_t)Mx * nk;
        a.tw = twy;
        int64_t cap = (int64_t)My * Mx * nk;
        check_strided(a, 1, &cap, &cap, 0);
        HIPCHK(launch_strided(My, prec, prec == PREC_FP64, prec == PREC_FP64, dir, a, 1, st));
    }
    void z(const void* in, void* out, int dir)
    {
        ZPlainArgs a{in, out, Mx, My, twz};
        int64_t np = (int64_t)Mx * My;
        int grid = (int)std::min<int64_t>((np + 1) / 2, 65536);
        HIPCHK(launch_zplain(Mz, prec, dir, a, grid, st));
    }
};

// ===========================================================================
// Fixed device scratch for spns_invariant_spectra_noalloc: the shell sums'
// block partials and results live here instead of in memory the library
// allocates.  Each GPU holds its own copy (4 MiB) until the process ends.
// In host code `scratch` names a host-side stand-in, never device memory:
// only the address from hipGetSymbolAddress (noalloc_scratch) is passed to
// kernels and copies.  Every call on one GPU shares that GPU's array, so
// calls that overlap on one GPU race.
// ===========================================================================
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 19;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, looked up once per device with
// that device current (spns_create calls it); nullptr if the lookup failed.
static double* noalloc_scratch(int device, hipError_t* err)
{
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? (double*)p : nullptr;
        if (g_scratch_err[device] != hipSuccess) (void)hipGetLastError();
        g_scratch_done[device] = true;
    }
    *err = g_scratch_err[device];
    return g_scratch_ptr[device];
}

// ===========================================================================
// Solver context
// ===========================================================================
struct spns_ctx {
    spns_config cfg{};
    int Nx, Ny, Nz, Mx, My, Mz, Nx_2, Ny_2, Nz_2;
    int nk1, nk2, nk3, nbins, last_rad, last_rad_max, prec, les, rot, sa;
    int ActualMx, ActualMy, ActualMz;
    int64_t nmodes;    // compact box (C ABI layout)
    int nk3p;          // k3 pitch of every device buffer (cache-line multiple)
    int64_t nmodes_p;  // nk2 * nk1 * nk3p
    double Ax, Ay, Az, delta, ct, nf;
    size_t tsz;  // bytes of one complex T
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device memory
    cx<double>* uk[3] = {};
    void* R[3] = {};  // R history, complex of the transform type (tsz bytes per mode)
    cx<double>* ukprev[3] = {};
    void* X1 = nullptr;
    void* X2 = nullptr;
    int64_t x1_slot = 0, x2_slot = 0;
    int n_x1 = 0, nfwd = 3;
    bool keep_prev = false;           // ukprev holds the start-of-step state
    double* invspec = nullptr;        // [5][nbins] invariant spectra, then the block partials
    double* scratch_dev = nullptr;    // this device's fixed scratch (spns_invariant_spectra_noalloc)
    hipError_t scratch_err = hipSuccess;
    void *twx = nullptr, *twy = nullptr, *twz = nullptr;
    DevScalars* scal = nullptr;
    double* hist_partial = nullptr;
    double* om_partial = nullptr;
    double* ed_partial = nullptr;  // [grid_epi] direct E sums of a reduce (QA energy check)
    double* skew_partial = nullptr;
    unsigned long long* mv = nullptr;  // [2]
    unsigned long long* mv_save = nullptr;  // [2] (the QA energy check keeps the CFL maxima)
    // x1v (device int): X1 slots 0-5 hold u_c and d/dy u_c of the current
    // state, written by the previous step's k_yepi_inv; the step's inverse y
    // pass then does nothing.  yinv: the step runs k_yepi_inv (grid_yinv).
    int* x1v = nullptr;
    bool yinv = false;
    int grid_yinv = 0;
    double* ek = nullptr;
    double* forced = nullptr;
    double* ft = nullptr;
    int grid_epi = 0, grid_z = 0;
    int grid_yepi = 0;  // > 0: the DNS step fuses its forward y pass with the epilogue (k_yepi)
    int zkernel = ZK_STAGED;  // z-pass kernel (environment SPNS_ZPASS=staged|nofuse|block|wave overrides)
    hipGraphExec_t graph = nullptr;
    bool graph_ready = false;
    std::string err;
    double last_ms = 0.0;

    char* xbase(void* p, int slot, int64_t slot_elems) const { return (char*)p + (size_t)slot * slot_elems * tsz; }
    void* X1f(int f) const { return xbase(X1, f, x1_slot); }
    void* X2f(int f) const { return xbase(X2, f, x2_slot); }
    void* Cf(int f) const { return xbase(X2, f, nmodes_p); }  // compact transformed fields, in X2's space

    void init(const spns_config& c)
    {
        cfg = c;
        if (cfg.Lx <= 0) cfg.Lx = 1;
        if (cfg.Ly <= 0) cfg.Ly = 1;
        if (cfg.Lz <= 0) cfg.Lz = 1;
        Nx = cfg.Nx; Ny = cfg.Ny; Nz = cfg.Nz;
        if (Nx < 2 || Ny < 2 || Nz < 2) throw Error(SPNS_EINVAL, "Nx, Ny, Nz must be >= 2");
        Mx = cfg.Mx > 0 ? cfg.Mx : (3 * Nx) / 2;
        My = cfg.My > 0 ? cfg.My : (3 * Ny) / 2;
        Mz = cfg.Mz > 0 ? cfg.Mz : (3 * Nz) / 2;
        cfg.Mx = Mx; cfg.My = My; cfg.Mz = Mz;
        for (int L : {Mx, My, Mz})
            if (!supported_length(L)) throw Error(SPNS_EUNSUPPORTED, "FFT length " + std::to_string(L) + " is not compiled in");
        prec = cfg.precision == SPNS_PREC_FP32_SPLIT ? PREC_FP32 : PREC_FP64;
        tsz = prec == PREC_FP64 ? 16 : 8;
        Nx_2 = Nx / 2; Ny_2 = Ny / 2; Nz_2 = Nz / 2;
        nk1 = 2 * Nx_2 + 1; nk2 = 2 * Ny_2 + 1; nk3 = Nz_2 + 1;
        if (Nx_2 >= (Mx + 1) / 2 || Ny_2 >= (My + 1) / 2 || Nz_2 >= Mz / 2)
            throw Error(SPNS_EINVAL, "retained modes do not fit the FFT grid");
        nmodes = (int64_t)nk2 * nk1 * nk3;
        nk3p = (nk3 + 15) / 16 * 16;
        nmodes_p = (int64_t)nk2 * nk1 * nk3p;
        // src/HIT_LES.h:31-58
        ActualMx = (Mx % 2 == 0) ? Mx / cfg.Lx : (Mx - 1) / cfg.Lx + 1;
        ActualMy = (My % 2 == 0) ? My / cfg.Ly : (My - 1) / cfg.Ly + 1;
        ActualMz = (Mz % 2 == 0) ? Mz / cfg.Lz : (Mz - 1) / cfg.Lz + 1;
        last_rad = std::min((Nx - 1) / 2, std::min((Ny - 1) / 2, (Nz - 1) / 2));
        last_rad_max = (int)std::ceil(std::sqrt(1.0 * Nx_2 * Nx_2 + 1.0 * Ny_2 * Ny_2 + 1.0 * Nz_2 * Nz_2));
        nbins = last_rad_max + 1;
        delta = cfg.delta > 0 ? cfg.delta
                              : 2.0 * kPi / std::cbrt((double)(cfg.Lx * cfg.Ly * cfg.Lz) * ActualMx * ActualMy * ActualMz);
        Ax = 2.0 * kPi / ActualMx / cfg.Lx;
        Ay = 2.0 * kPi / ActualMy / cfg.Ly;
        Az = 2.0 * kPi / ActualMz / cfg.Lz;
        ct = cfg.C_Smag * cfg.C_Smag * delta * delta;  // src/HIT_LES.cpp:132
        nf = (double)((int64_t)Mx * My * Mz);
        les = cfg.C_Smag > 0.0;
        rot = (std::fabs(cfg.omega_x) > 0 || std::fabs(cfg.omega_y) > 0 || std::fabs(cfg.omega_z) > 0);
        sa = cfg.isSelfAdaptive != 0;
        nfwd = les ? 9 : 3;
        keep_prev = sa || cfg.gradient_invariants || cfg.keep_step_start;

        zkernel = cfg.zpass_variant == 1 ? ZK_BLOCK
                  : cfg.zpass_variant == 2 ? ZK_WAVE
                  : cfg.zpass_variant == 3 ? ZK_STAGED_NOFUSE
                                           : ZK_STAGED;
        if (const char* zk = std::getenv("SPNS_ZPASS")) {
            const std::string v(zk);
            zkernel = v == "wave" ? ZK_WAVE : v == "block" ? ZK_BLOCK : v == "nofuse" ? ZK_STAGED_NOFUSE : ZK_STAGED;
        }
        HIPCHK(hipSetDevice(cfg.device));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        const size_t sb = sizeof(cx<double>) * (size_t)nmodes_p;
        for (int c = 0; c < 3; c++) {
            alloc((void**)&uk[c], sb);
            alloc(&R[c], tsz * (size_t)nmodes_p);
            if (keep_prev) alloc((void**)&ukprev[c], sb);
        }
        x1_slot = (int64_t)My * nk1 * nk3p;
        x2_slot = (int64_t)My * Mx * nk3p;
        n_x1 = std::max(6, nfwd);
        alloc(&X1, (size_t)n_x1 * x1_slot * tsz);
        alloc(&X2, (size_t)9 * x2_slot * tsz);
        twx = upload_twiddles(Mx, prec);
        twy = upload_twiddles(My, prec);
        twz = upload_twiddles(Mz, prec);
        alloc((void**)&scal, sizeof(DevScalars));
        grid_epi = (int)std::min<int64_t>((nmodes_p + 255) / 256, 1024);
        const int64_t npen = (int64_t)My * Mx;
        {
            ZArgs zq{};
            zq.kernel = zkernel;
            zq.pitch = nk3p;
            zq.nk3 = nk3;
            zq.les = les;
            grid_z = zpass_grid(Mz, prec, zq, npen);
            if (grid_z <= 0) throw Error(SPNS_EHIP, "z pass: no launch configuration for this length");
        }
        alloc((void**)&hist_partial, sizeof(double) * (size_t)grid_epi * nbins);
        alloc((void**)&om_partial, sizeof(double) * grid_epi);
        alloc((void**)&ed_partial, sizeof(double) * grid_epi);
        alloc((void**)&skew_partial, sizeof(double) * 2 * grid_z);
        alloc((void**)&mv, sizeof(unsigned long long) * 2);
        alloc((void**)&mv_save, sizeof(unsigned long long) * 2);
        alloc((void**)&x1v, sizeof(int));
        alloc((void**)&ek, sizeof(double) * nbins);
        alloc((void**)&ft, sizeof(double) * (nbins + 1));
        HIPCHK(hipMemsetAsync(skew_partial, 0, sizeof(double) * 2 * grid_z, st));
        HIPCHK(hipMemsetAsync(X1, 0, (size_t)n_x1 * x1_slot * tsz, st));
        HIPCHK(hipMemsetAsync(X2, 0, (size_t)9 * x2_slot * tsz, st));
        HIPCHK(hipMemsetAsync(mv, 0, sizeof(unsigned long long) * 2, st));
        HIPCHK(hipMemsetAsync(x1v, 0, sizeof(int), st));
        for (int c = 0; c < 3; c++) {
            HIPCHK(hipMemsetAsync(uk[c], 0, sb, st));
            HIPCHK(hipMemsetAsync(R[c], 0, tsz * (size_t)nmodes_p, st));
            if (keep_prev) HIPCHK(hipMemsetAsync(ukprev[c], 0, sb, st));
        }
        // fused forward y pass + epilogue for DNS steps (environment
        // SPNS_EPI=split keeps the two kernels apart)
        {
            const char* ev = std::getenv("SPNS_EPI");
            const bool split = ev && std::string(ev) == "split";
            grid_yepi = 0;
            if (!les && !split && fwd_y_args(3).padded_ok) {
                const int tb = strided_tile(My, prec);
                const int64_t items = tb > 0 ? (int64_t)nk1 * ((nk3 + tb - 1) / tb) : 0;
                if (items > 0) grid_yepi = yepi_grid(My, prec, nbins, nk2, items, grid_epi);
                // the next step's inverse y pass inside the fused kernel
                // (environment SPNS_YINV=0 keeps it a separate pass)
                const char* yv = std::getenv("SPNS_YINV");
                yinv = false;
                grid_yinv = 0;
                if (grid_yepi > 0 && !cfg.isNotCubic && !(yv && yv[0] == '0') && inv_y_args(false).padded_ok) {
                    grid_yinv = yepi_grid(My, prec, nbins, nk2, items, grid_epi, true);
                    yinv = grid_yinv > 0;
                }
            }
        }
        DevScalars s{};
        s.nu = cfg.nu;
        s.kappa = 0.5;  // src/HIT_LES.cpp:1053
        HIPCHK(hipMemcpyAsync(scal, &s, sizeof(s), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    // Takes the physics of c (viscosity, C_s, CFL constant, rotation, fixed
    // step, forcing, filter width) for a run of the same shape class and
    // clears state, history, time and iteration count; the device buffers
    // stay.  The caller sets an initial field and calls spns_finish_init.
    void reconfigure(const spns_config& c)
    {
        const int cLx = c.Lx > 0 ? c.Lx : 1, cLy = c.Ly > 0 ? c.Ly : 1, cLz = c.Lz > 0 ? c.Lz : 1;
        const int cMx = c.Mx > 0 ? c.Mx : (3 * c.Nx) / 2, cMy = c.My > 0 ? c.My : (3 * c.Ny) / 2,
                  cMz = c.Mz > 0 ? c.Mz : (3 * c.Nz) / 2;
        const bool same = c.Nx == Nx && c.Ny == Ny && c.Nz == Nz && cMx == Mx && cMy == My && cMz == Mz &&
                          c.precision == cfg.precision && (c.C_Smag > 0.0) == les && cLx == cfg.Lx && cLy == cfg.Ly &&
                          cLz == cfg.Lz && c.isNotCubic == cfg.isNotCubic && c.isSelfAdaptive == cfg.isSelfAdaptive &&
                          c.gradient_invariants == cfg.gradient_invariants && c.keep_step_start == cfg.keep_step_start &&
                          c.zpass_variant == cfg.zpass_variant && c.device == cfg.device;
        if (!same) throw Error(SPNS_EINVAL, "reconfigure: the grid, precision, model and buffers must stay the same");
        HIPCHK(hipStreamSynchronize(st));
        const spns_config old = cfg;
        cfg = c;
        cfg.Lx = cLx; cfg.Ly = cLy; cfg.Lz = cLz;
        cfg.Mx = Mx; cfg.My = My; cfg.Mz = Mz;
        delta = cfg.delta > 0 ? cfg.delta
                              : 2.0 * kPi / std::cbrt((double)(cfg.Lx * cfg.Ly * cfg.Lz) * ActualMx * ActualMy * ActualMz);
        ct = cfg.C_Smag * cfg.C_Smag * delta * delta;  // src/HIT_LES.cpp:132
        rot = (std::fabs(cfg.omega_x) > 0 || std::fabs(cfg.omega_y) > 0 || std::fabs(cfg.omega_z) > 0);
        (void)old;
        if (graph) {
            hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        graph_ready = false;
        const size_t sb = sizeof(cx<double>) * (size_t)nmodes_p;
        for (int c3 = 0; c3 < 3; c3++) {
            HIPCHK(hipMemsetAsync(uk[c3], 0, sb, st));
            HIPCHK(hipMemsetAsync(R[c3], 0, tsz * (size_t)nmodes_p, st));
            if (keep_prev) HIPCHK(hipMemsetAsync(ukprev[c3], 0, sb, st));
        }
        DevScalars s{};
        s.nu = cfg.nu;
        s.kappa = 0.5;  // src/HIT_LES.cpp:1053
        write_scalars(s);
        HIPCHK(hipStreamSynchronize(st));
    }

    void alloc(void** p, size_t bytes)
    {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            *p = nullptr;
            throw Error(SPNS_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
    }

    void release()
    {
        if (graph) hipGraphExecDestroy(graph);
        for (int c = 0; c < 3; c++) { hipFree(uk[c]); hipFree(R[c]); hipFree(ukprev[c]); }
        hipFree(X1); hipFree(X2); hipFree(twx); hipFree(twy); hipFree(twz);
        hipFree(scal); hipFree(hist_partial); hipFree(om_partial); hipFree(ed_partial); hipFree(skew_partial); hipFree(mv); hipFree(mv_save);
        hipFree(x1v);
        hipFree(ek); hipFree(forced); hipFree(ft); hipFree(invspec);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (st) hipStreamDestroy(st);
    }

    // ---------------------------------------------------------------- passes
    // y passes: rows = y (or i2), batch = (i1 outer, i3 contiguous)
    // Each pass's argument struct (with the outer extent A set) is built by
    // its *_args method, so the batched step (spns_batch) launches exactly
    // the kernels the single-run step does, on the same structs.
    StridedArgs inv_y_args(bool velocity_only, cx<double>* const* src = nullptr)
    {
        if (!src) src = uk;
        StridedArgs a{};
        int n = 0;
        for (int c = 0; c < 3; c++)  // u_c, and d/dy u_c from the same read of u_c
            a.job[n++] = StridedJob{src[c], X1f(c), 0, 0, velocity_only ? nullptr : X1f(3 + c)};
        a.njobs = n;
        a.B = nk3; a.nin = nk2; a.nout = My;
        a.in_row_stride = (int64_t)nk1 * nk3p; a.out_row_stride = (int64_t)nk1 * nk3p;
        a.in_a_stride = nk3p; a.out_a_stride = nk3p;
        a.tw = twy;
        int64_t icap[9], ocap[9];
        for (int j = 0; j < n; j++) { icap[j] = nmodes_p; ocap[j] = x1_slot; }
        check_strided(a, nk1, icap, ocap, strided_tile(My, prec));
        a.A = nk1;
        return a;
    }
    // The step's inverse y pass (all six fields of the current state) does
    // nothing while x1v is set; any other use rewrites X1 and clears it.
    void inv_y(bool velocity_only, cx<double>* const* src = nullptr)
    {
        StridedArgs a = inv_y_args(velocity_only, src);
        if (yinv && !velocity_only && !src) a.skip = x1v;
        else x1_stale();
        HIPCHK(launch_strided(My, prec, 1, prec == PREC_FP64, +1, a, a.A, st));
    }

    // X1 no longer holds the current state's inverse y pass (the state
    // changed, or X1 was rewritten outside a step)
    void x1_stale()
    {
        if (yinv) HIPCHK(launch_set_int(x1v, 0, st));
    }

    // x passes: rows = x (or i1), batch = (y outer, i3 contiguous)
    StridedArgs inv_x_args(bool velocity_only)
    {
        StridedArgs a{};
        int n = 0;
        for (int c = 0; c < 3; c++) {  // u_c and d/dx u_c from one read, then d/dy u_c
            a.job[n++] = StridedJob{X1f(c), X2f(c), 0, 0, velocity_only ? nullptr : X2f(3 + c)};
            if (!velocity_only) a.job[n++] = StridedJob{X1f(3 + c), X2f(6 + c), 0, 0};
        }
        a.njobs = n;
        a.B = nk3; a.nin = nk1; a.nout = Mx;
        a.in_row_stride = nk3p; a.in_a_stride = (int64_t)nk1 * nk3p;
        a.out_row_stride = nk3p; a.out_a_stride = (int64_t)Mx * nk3p;
        a.tw = twx;
        int64_t icap[9], ocap[9];
        for (int j = 0; j < n; j++) { icap[j] = x1_slot; ocap[j] = x2_slot; }
        check_strided(a, My, icap, ocap, strided_tile(Mx, prec));
        a.A = My;
        return a;
    }
    void inv_x(bool velocity_only)
    {
        const StridedArgs a = inv_x_args(velocity_only);
        HIPCHK(launch_strided(Mx, prec, prec == PREC_FP64, prec == PREC_FP64, +1, a, a.A, st));
    }

    ZArgs zpass_args(void* phys)
    {
        ZArgs z{};
        for (int f = 0; f < 9; f++) z.f[f] = X2f(f);
        z.nk3 = nk3;
        z.pitch = nk3p;
        z.les = les;
        z.sa = sa;
        z.kernel = zkernel;
        z.npencil = (int64_t)My * Mx;
        z.Ax = Ax; z.Ay = Ay; z.Az = Az;
        z.ct = ct;
        z.maxvel_bits = mv;
        z.maxvel_sa_bits = mv + 1;
        z.skew_partial = skew_partial;
        z.phys = phys;
        z.tw = twz;
        if (z.npencil * nk3p > x2_slot || nk3 > nk3p) throw Error(SPNS_EINVAL, "z pass exceeds its buffer");
        return z;
    }
    void zpass(int mode, void* phys)
    {
        ZArgs z = zpass_args(phys);
        if (mode == 2) {
            z.kernel = ZK_STAGED_NOFUSE;
            const hipError_t e = launch_zpass(Mz, prec, mode, z, grid_z, st);
            if (e == hipErrorNotSupported)
                throw Error(SPNS_EUNSUPPORTED, "gradient invariants: no staged z pass at this length and precision");
            HIPCHK(e);
            return;
        }
        HIPCHK(launch_zpass(Mz, prec, mode, z, grid_z, st));
    }

    // Shell spectra of the velocity-gradient invariants QG QS RG RS V2 of the
    // field at the start of the last step (src/HIT_LES.cpp:1297-1319: the
    // reference evaluates the gradient arrays its last convective term left,
    // i.e. of that field; before any step, of the current field).  Runs the
    // inverse passes on that field, the z pass in its invariants mode (five
    // fields, forward along z), the forward x and y passes on five fields and
    // the shell sums.  out: [5][nbins].
    // block count of the shell sums of k_inv_spectra
    int inv_spectra_grid() const { return std::min(1024, nk1 * nk2); }
    // device doubles the sums need: [5][nbins] results + [grid][5][nbins] partials
    size_t inv_spectra_scratch() const { return ((size_t)inv_spectra_grid() + 1) * 5 * nbins; }

    // Runs the passes and the shell sums; the [5][nbins] result lands in
    // work[0 .. 5 nbins), the block partials after it (inv_spectra_scratch()
    // doubles in all, fully overwritten).
    void invariant_spectra_dev(double* work)
    {
        if (!cfg.gradient_invariants)
            throw Error(SPNS_EUNSUPPORTED, "invariant spectra need spns_config.gradient_invariants at creation");
        const DevScalars sc = read_scalars();
        cx<double>* const* src = (sc.iter > 0) ? ukprev : uk;
        inv_y(false, src);
        inv_x(false);
        zpass(2, nullptr);
        fwd_x(5);
        fwd_y(5);
        InvSpecArgs a{};
        for (int f = 0; f < 5; f++) a.f[f] = Cf(f);
        a.nk1 = nk1; a.nk2 = nk2; a.nk3 = nk3; a.nk3p = nk3p; a.nbins = nbins;
        a.notcubic = cfg.isNotCubic; a.Lx = cfg.Lx; a.Ly = cfg.Ly; a.Lz = cfg.Lz; a.Mz = Mz;
        a.inv_nf2 = 1.0 / (nf * nf);
        a.partial = work + (size_t)5 * nbins;
        HIPCHK(launch_inv_spectra(prec, a, inv_spectra_grid(), work, st));
    }

    void invariant_spectra(double* out)
    {
        if (!invspec) alloc((void**)&invspec, sizeof(double) * inv_spectra_scratch());
        invariant_spectra_dev(invspec);
        HIPCHK(hipMemcpyAsync(out, invspec, sizeof(double) * 5 * nbins, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    StridedArgs fwd_x_args(int nfields = 0)
    {
        const int nf = nfields > 0 ? nfields : nfwd;
        StridedArgs a{};
        for (int f = 0; f < nf; f++) a.job[f] = StridedJob{X2f(f), X1f(f), 0, 0};
        a.njobs = nf;
        a.B = nk3; a.nin = Mx; a.nout = nk1;
        a.in_row_stride = nk3p; a.in_a_stride = (int64_t)Mx * nk3p;
        a.out_row_stride = nk3p; a.out_a_stride = (int64_t)nk1 * nk3p;
        a.tw = twx;
        int64_t icap[9], ocap[9];
        for (int j = 0; j < nf; j++) { icap[j] = x2_slot; ocap[j] = x1_slot; }
        check_strided(a, My, icap, ocap, strided_tile(Mx, prec));
        a.A = My;
        return a;
    }
    void fwd_x(int nfields = 0)
    {
        const StridedArgs a = fwd_x_args(nfields);
        HIPCHK(launch_strided(Mx, prec, prec == PREC_FP64, prec == PREC_FP64, -1, a, a.A, st));
    }

    StridedArgs fwd_y_args(int nf)
    {
        StridedArgs a{};
        for (int f = 0; f < nf; f++) a.job[f] = StridedJob{X1f(f), Cf(f), 0, 0};
        a.njobs = nf;
        a.B = nk3; a.nin = My; a.nout = nk2;
        a.in_row_stride = (int64_t)nk1 * nk3p; a.out_row_stride = (int64_t)nk1 * nk3p;
        a.in_a_stride = nk3p; a.out_a_stride = nk3p;
        a.tw = twy;
        int64_t icap[9], ocap[9];
        for (int j = 0; j < nf; j++) { icap[j] = x1_slot; ocap[j] = nmodes_p; }
        check_strided(a, nk1, icap, ocap, strided_tile(My, prec));
        if ((int64_t)nf * nmodes_p > 9 * x2_slot) throw Error(SPNS_EINVAL, "compact fields exceed X2");
        a.A = nk1;
        return a;
    }

    void fwd_y(int nfields = 0)
    {
        const StridedArgs a = fwd_y_args(nfields > 0 ? nfields : nfwd);
        HIPCHK(launch_strided(My, prec, prec == PREC_FP64, prec == PREC_FP64, -1, a, a.A, st));
    }

    // the DNS step's forward y pass fused with its epilogue (k_yepi) when
    // grid_yepi > 0, else fwd_y then k_epilogue; returns the block count
    // of the shell partials the finalize sums
    int fwd_y_epilogue()
    {
        if (yinv) {
            const StridedArgs ia = inv_y_args(false);
            HIPCHK(launch_yepi(My, prec, fwd_y_args(3), nk1, epi_args(EPI_STEP), grid_yinv, st, nullptr, &ia, x1v));
            return grid_yinv;
        }
        if (grid_yepi > 0) {
            HIPCHK(launch_yepi(My, prec, fwd_y_args(3), nk1, epi_args(EPI_STEP), grid_yepi, st));
            return grid_yepi;
        }
        fwd_y();
        HIPCHK(launch_epilogue(prec, epi_args(EPI_STEP), grid_epi, st));
        return grid_epi;
    }

    EpiArgs epi_args(int mode)
    {
        EpiArgs e{};
        for (int f = 0; f < nfwd; f++) e.conv[f] = Cf(f);
        for (int c = 0; c < 3; c++) { e.uk[c] = uk[c]; e.R[c] = R[c]; e.ukprev[c] = ukprev[c]; }
        e.nk1 = nk1; e.nk2 = nk2; e.nk3 = nk3; e.nk3p = nk3p; e.nbins = nbins;
        e.nmodes = nmodes_p;
        e.notcubic = cfg.isNotCubic; e.Lx = cfg.Lx; e.Ly = cfg.Ly; e.Lz = cfg.Lz; e.Mz = Mz;
        e.mode = mode; e.les = les; e.rot = rot; e.force_on = cfg.force_on;
        e.kfmin = cfg.force_kmin; e.kfmax = cfg.force_kmax; e.sa = sa;
        e.keepprev = keep_prev && !sa;
        e.nf = nf;
        e.ox = cfg.omega_x; e.oy = cfg.omega_y; e.oz = cfg.omega_z;
        e.scal = scal;
        e.hist_partial = hist_partial;
        e.om_partial = om_partial;
        e.ed_partial = mode == EPI_REDUCE ? ed_partial : nullptr;
        return e;
    }

    FinArgs fin_args(int flags)
    {
        FinArgs f{};
        f.scal = scal;
        f.flags = flags;
        f.nbins = nbins; f.nblk_epi = grid_epi; f.nblk_z = grid_z;
        f.hist_partial = hist_partial; f.om_partial = om_partial; f.ed_partial = ed_partial; f.skew_partial = skew_partial;
        f.maxvel_bits = mv; f.maxvel_sa_bits = mv + 1;
        f.ek = ek;
        f.kfmin = cfg.force_kmin; f.kfmax = cfg.force_kmax; f.force_eps = cfg.force_on ? cfg.force_eps : 0.0;
        f.Ax = Ax; f.Ay = Ay; f.Az = Az; f.C_At = cfg.C_At; f.fixed_dt = cfg.fixed_dt;
        f.sa = sa;
        // self-adaptive constants (src/HIT_LES.cpp:1049-1078)
        f.Ah = (Ax + Ay + Az) / 3.0;
        f.t1 = 0.9302468;
        f.kk[0] = 0.73782212; f.kk[1] = 0.44660387;
        f.phi[0] = std::atan(164.0 / 99.0); f.phi[1] = kPi / 3.0; f.phi[2] = kPi * 3.0 * 3.0 / (5.0 * 5.0);
        const double c[14] = {0.0647998, -0.386022, 3.72945, -9.38143, 7.06574, 2403400.0, -5018490.0,
                              2620140.0, 2945.0, -6665.76, 3790.54, 4.80513, -16.9473, 15.0155};
        std::memcpy(f.c, c, sizeof(c));
        return f;
    }

    void enqueue_step()
    {
        HIPCHK(launch_clear(mv, mv + 1, st));
        inv_y(false);
        inv_x(false);
        zpass(0, nullptr);
        fwd_x();
        FinArgs f = fin_args(FIN_REDUCE | FIN_SKEW | FIN_MAXVEL | FIN_DT | FIN_ADVANCE);
        f.nblk_epi = fwd_y_epilogue();
        HIPCHK(launch_finalize(f, st));
    }

    void reduce_state()
    {
        HIPCHK(launch_epilogue(prec, epi_args(EPI_REDUCE), grid_epi, st));
        HIPCHK(launch_finalize(fin_args(FIN_REDUCE | FIN_EFOUR), st));
    }

    // HIT::Check_PhysFour_Energy (src/HIT_LES.cpp:1216-1252) on the current
    // state: E summed directly over the retained modes (EkFour), by shells
    // (EkSpNS) and in physical space (EkReal, 1/2 sum u.u / (Mx My Mz) from
    // a velocity pass).  The step's scalars, spectrum and CFL maxima are
    // left as they were.
    void check_phys_four_energy(double* ek_real, double* ek_four, double* ek_spns)
    {
        const DevScalars before = read_scalars();
        std::vector<double> bins(nbins);
        HIPCHK(hipMemcpyAsync(bins.data(), ek, sizeof(double) * nbins, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(mv_save, mv, sizeof(unsigned long long) * 2, hipMemcpyDeviceToDevice, st));
        reduce_state();
        velocity_pass(nullptr);
        const DevScalars after = read_scalars();
        *ek_real = 0.5 * after.ek_phys / nf;  // src/HIT_LES.cpp:1236-1240
        *ek_four = after.ek_four;
        *ek_spns = after.ek;
        write_scalars(before);
        HIPCHK(hipMemcpyAsync(ek, bins.data(), sizeof(double) * nbins, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(mv, mv_save, sizeof(unsigned long long) * 2, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    // u,v,w from the state: CFL maxima (and the physical field if phys)
    // Physical velocity of the state at the start of the last step (the
    // current state before any step) into phys.  The maxima the z pass
    // accumulates stay in mv, which every step and velocity pass clears and
    // only a finalize with FIN_MAXVEL reads, so the CFL state is unchanged.
    void velocity_step_start(void* phys)
    {
        const DevScalars sc = read_scalars();
        if (sc.iter > 0 && !keep_prev)
            throw Error(SPNS_EUNSUPPORTED, "the step-start field needs spns_config.keep_step_start at creation");
        cx<double>* const* src = (sc.iter > 0) ? ukprev : uk;
        inv_y(true, src);
        inv_x(true);
        zpass(1, phys);
    }

    void velocity_pass(void* phys)
    {
        HIPCHK(launch_clear(mv, mv + 1, st));
        inv_y(true);
        inv_x(true);
        zpass(1, phys);
        HIPCHK(launch_finalize(fin_args(FIN_MAXVEL | FIN_EPHYS), st));
    }

    void step(int n)
    {
        HIPCHK(hipEventRecord(ev0, st));
        if (cfg.use_graph) {
            if (!graph_ready) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_step();
                } catch (...) {
                    hipGraph_t dummy;
                    hipStreamEndCapture(st, &dummy);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(st, &g));
                HIPCHK(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
                HIPCHK(hipGraphDestroy(g));
                graph_ready = true;
            }
            for (int i = 0; i < n; i++) HIPCHK(hipGraphLaunch(graph, st));
        } else {
            for (int i = 0; i < n; i++) enqueue_step();
        }
        HIPCHK(hipEventRecord(ev1, st));
    }

    DevScalars read_scalars()
    {
        DevScalars s;
        HIPCHK(hipMemcpyAsync(&s, scal, sizeof(s), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return s;
    }

    void write_scalars(const DevScalars& s)
    {
        HIPCHK(hipMemcpyAsync(scal, &s, sizeof(s), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    void upload_state(const std::vector<std::complex<double>>& h)
    {
        // h: [3][nmodes]; modes masked by the length factors are zeroed
        std::vector<std::complex<double>> m(h);
        if (cfg.isNotCubic) {
            for (int64_t q = 0; q < nmodes; q++) {
                int i3 = (int)(q % nk3);
                int64_t r = q / nk3;
                int i1 = (int)(r % nk1), i2 = (int)(r / nk1);
                int k1 = i1 <= nk1 / 2 ? i1 : i1 - nk1, k2 = i2 <= nk2 / 2 ? i2 : i2 - nk2, k3 = i3;
                if ((k1 % cfg.Lx) || (k2 % cfg.Ly) || (k3 % cfg.Lz))
                    for (int c = 0; c < 3; c++) m[c * nmodes + q] = 0.0;
            }
        }
        for (int c = 0; c < 3; c++)
            HIPCHK(hipMemcpy2DAsync(uk[c], sizeof(cx<double>) * nk3p, m.data() + c * nmodes, sizeof(cx<double>) * nk3,
                                    sizeof(cx<double>) * nk3, (size_t)nk2 * nk1, hipMemcpyHostToDevice, st));
        reset_history();
    }

    void reset_history()
    {
        for (int c = 0; c < 3; c++) {
            HIPCHK(hipMemsetAsync(R[c], 0, tsz * (size_t)nmodes_p, st));
            if (keep_prev) HIPCHK(hipMemsetAsync(ukprev[c], 0, sizeof(cx<double>) * nmodes_p, st));
        }
        DevScalars s = read_scalars();
        s.time = 0.0;
        s.iter = 0;
        write_scalars(s);
    }

    std::vector<std::complex<double>> download_state(cx<double>* const* src)
    {
        std::vector<std::complex<double>> h(3 * (size_t)nmodes);
        for (int c = 0; c < 3; c++)
            HIPCHK(hipMemcpy2DAsync(h.data() + c * nmodes, sizeof(cx<double>) * nk3, src[c], sizeof(cx<double>) * nk3p,
                                    sizeof(cx<double>) * nk3, (size_t)nk2 * nk1, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return h;
    }

    int64_t compact_index(int k1, int k2, int k3) const
    {
        int i1 = k1 >= 0 ? k1 : nk1 + k1, i2 = k2 >= 0 ? k2 : nk2 + k2;
        return ((int64_t)i2 * nk1 + i1) * nk3 + k3;
    }
    bool retained(int k1, int k2, int k3) const
    {
        if (std::abs(k1) > Nx_2 || std::abs(k2) > Ny_2 || k3 > Nz_2 || k3 < 0) return false;
        if (cfg.isNotCubic && ((k1 % cfg.Lx) || (k2 % cfg.Ly) || (k3 % cfg.Lz))) return false;
        return true;
    }

    void forcing_energy_cascade(const double* fek, int64_t last_input_rad, int nullify)
    {
        reduce_state();
        const int last_eff = (int)std::min<int64_t>(last_input_rad, last_rad);
        if (last_eff < 0) return;
        std::vector<double> f(fek, fek + last_eff + 1);
        if (forced) hipFree(forced);
        forced = nullptr;
        alloc((void**)&forced, sizeof(double) * (last_eff + 1));
        HIPCHK(hipMemcpyAsync(forced, f.data(), sizeof(double) * (last_eff + 1), hipMemcpyHostToDevice, st));
        ShellArgs s = shell_args();
        s.ft = ft;
        s.last_eff = last_eff;
        s.nullify = nullify;
        if (last_eff + 1 > nbins + 1) throw Error(SPNS_EINVAL, "forcing shells exceed the spectrum");
        HIPCHK(launch_forcing(ek, forced, ft, last_eff, s, st));
        if (cfg.isComplexConjugateCorrected) HIPCHK(launch_ccc(shell_args(), st));
        reduce_state();
        HIPCHK(hipStreamSynchronize(st));
    }

    ShellArgs shell_args()
    {
        ShellArgs s{};
        for (int c = 0; c < 3; c++) s.uk[c] = uk[c];
        s.nk1 = nk1; s.nk2 = nk2; s.nk3 = nk3; s.nk3p = nk3p; s.nmodes = nmodes_p;
        s.notcubic = cfg.isNotCubic; s.Lx = cfg.Lx; s.Ly = cfg.Ly; s.Lz = cfg.Lz;
        return s;
    }
};

// ===========================================================================
// Batched stepping of several runs of one shape class (spns_batch_*).
// Each slot is an ordinary context (its own buffers, created and initialised
// as usual).  A batched step launches every pass once over all slots: the
// kernels take the slots' argument structs from device arrays (slot =
// blockIdx.y), built by the same *_args methods the single-run step uses,
// with the single run's grids, so every slot computes exactly what it would
// alone.  Per-slot steps-left counters let slots advance different step
// counts within one captured chunk; a slot at 0 skips every launch.
// ===========================================================================
struct spns_batch {
    int device = 0, nslots = 0, chunk = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;     // after the last queued work (chunk, uploads or fetches)
    hipEvent_t ev_sc = nullptr;  // after the last copy of the slots' scalars (hsc)
    std::vector<spns_ctx*> slot;
    spns_ctx* rep = nullptr;  // a slot context giving the class's shapes
    // per-slot argument structs: pinned host copies (so their uploads are
    // asynchronous on the batch's stream) and device arrays
    struct SlotArgs {
        StridedArgs invy, invx, fwdx, fwdy;
        StridedArgs invy2;  // the next step's inverse y pass inside the fused kernel (yinv classes)
        ZArgs z;
        EpiArgs epi;
        FinArgs fin;
        ClearArgs clr;
    };
    SlotArgs* hs = nullptr;
    // pinned copies of each slot's device scalars, refreshed after every
    // chunk on the batch's stream: the loop reads a run's status without a
    // synchronous device read (which would wait behind other streams' work
    // sharing its hardware queue)
    DevScalars* hsc = nullptr;
    StridedArgs *d_invy = nullptr, *d_invx = nullptr, *d_fwdx = nullptr, *d_fwdy = nullptr, *d_invy2 = nullptr;
    ZArgs* d_z = nullptr;
    EpiArgs* d_epi = nullptr;
    FinArgs* d_fin = nullptr;
    ClearArgs* d_clr = nullptr;
    int* d_left = nullptr;
    hipGraphExec_t graph = nullptr;
    std::string err;

    void init(int dev, int n, int ch)
    {
        if (n < 1 || n > kMaxBatchSlots) throw Error(SPNS_EINVAL, "nslots must be in 1.." + std::to_string(kMaxBatchSlots));
        if (ch < 1) throw Error(SPNS_EINVAL, "chunk must be >= 1");
        device = dev; nslots = n; chunk = ch;
        slot.assign(n, nullptr);
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipHostMalloc((void**)&hs, sizeof(SlotArgs) * n, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&hsc, sizeof(DevScalars) * n, hipHostMallocDefault));
        std::memset((void*)hs, 0, sizeof(SlotArgs) * n);
        std::memset((void*)hsc, 0, sizeof(DevScalars) * n);
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_sc, hipEventDisableTiming));
        auto al = [&](void** p, size_t bytes) {
            hipError_t e = hipMalloc(p, bytes);
            if (e != hipSuccess) { *p = nullptr; throw Error(SPNS_ENOMEM, "batch argument arrays: " + std::string(hipGetErrorString(e))); }
        };
        al((void**)&d_invy, sizeof(StridedArgs) * n); al((void**)&d_invx, sizeof(StridedArgs) * n);
        al((void**)&d_invy2, sizeof(StridedArgs) * n);
        al((void**)&d_fwdx, sizeof(StridedArgs) * n); al((void**)&d_fwdy, sizeof(StridedArgs) * n);
        al((void**)&d_z, sizeof(ZArgs) * n); al((void**)&d_epi, sizeof(EpiArgs) * n);
        al((void**)&d_fin, sizeof(FinArgs) * n); al((void**)&d_clr, sizeof(ClearArgs) * n);
        al((void**)&d_left, sizeof(int) * n);
        SlotSteps z{};
        HIPCHK(launch_set_left(d_left, z, nslots, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    void release()
    {
        if (st) hipStreamSynchronize(st);
        if (graph) hipGraphExecDestroy(graph);
        hipFree(d_invy); hipFree(d_invx); hipFree(d_fwdx); hipFree(d_fwdy); hipFree(d_invy2); hipFree(d_z); hipFree(d_epi);
        hipFree(d_fin); hipFree(d_clr); hipFree(d_left);
        if (hs) hipHostFree(hs);
        if (hsc) hipHostFree(hsc);
        if (ev) hipEventDestroy(ev);
        if (ev_sc) hipEventDestroy(ev_sc);
        if (st) hipStreamDestroy(st);
    }

    // a context joins the class if its shapes, model and kernel choices match
    // the representative's
    static bool same_class(const spns_ctx* a, const spns_ctx* b)
    {
        return a->Nx == b->Nx && a->Ny == b->Ny && a->Nz == b->Nz && a->Mx == b->Mx && a->My == b->My && a->Mz == b->Mz &&
               a->prec == b->prec && a->les == b->les && a->nfwd == b->nfwd && a->zkernel == b->zkernel &&
               a->grid_z == b->grid_z && a->grid_epi == b->grid_epi && a->grid_yepi == b->grid_yepi && a->nbins == b->nbins &&
               a->yinv == b->yinv && a->grid_yinv == b->grid_yinv && a->cfg.device == b->cfg.device;
    }

    void attach(int s, spns_ctx* c)
    {
        if (s < 0 || s >= nslots) throw Error(SPNS_EINVAL, "slot out of range");
        if (!c) throw Error(SPNS_EINVAL, "null context");
        if (c->cfg.device != device) throw Error(SPNS_EINVAL, "context on another device");
        spns_ctx* r = rep;
        if (!r)
            for (spns_ctx* o : slot) if (o) { r = o; break; }
        if (r && !same_class(r, c)) throw Error(SPNS_EINVAL, "context does not match the batch's shape class");
        if (!batch_length(c->Mx) || !batch_length(c->My) || !batch_length(c->Mz))
            throw Error(SPNS_EUNSUPPORTED, "no batched kernels for this grid's FFT lengths");
        HIPCHK(hipStreamSynchronize(st));        // no copy in flight reads this slot's pinned structs
        c->x1_stale();                           // the first batched step runs the inverse y pass
        HIPCHK(hipStreamSynchronize(c->st));     // the context's own work is done
        SlotArgs& a = hs[s];
        a.invy = c->inv_y_args(false);
        a.invy2 = StridedArgs{};
        if (c->yinv) {
            // as in the single-run step: the fused kernel runs the next step's
            // inverse y pass and sets the slot's flag, which the batched inverse
            // y pass reads
            a.invy.skip = c->x1v;
            a.invy2 = c->inv_y_args(false);
            a.invy2.setflag = c->x1v;
        }
        a.invx = c->inv_x_args(false);
        a.z = c->zpass_args(nullptr);
        a.fwdx = c->fwd_x_args();
        a.fwdy = c->fwd_y_args(c->nfwd);
        a.epi = c->epi_args(EPI_STEP);
        a.fin = c->fin_args(FIN_REDUCE | FIN_SKEW | FIN_MAXVEL | FIN_DT | FIN_ADVANCE);
        a.fin.nblk_epi = c->yinv ? c->grid_yinv : c->grid_yepi > 0 ? c->grid_yepi : c->grid_epi;
        a.clr = ClearArgs{c->mv, c->mv + 1};
        auto up = [&](void* d, const void* h, size_t bytes) { HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st)); };
        up(d_invy + s, &a.invy, sizeof(StridedArgs)); up(d_invx + s, &a.invx, sizeof(StridedArgs));
        up(d_z + s, &a.z, sizeof(ZArgs)); up(d_fwdx + s, &a.fwdx, sizeof(StridedArgs));
        up(d_fwdy + s, &a.fwdy, sizeof(StridedArgs)); up(d_epi + s, &a.epi, sizeof(EpiArgs));
        up(d_fin + s, &a.fin, sizeof(FinArgs)); up(d_clr + s, &a.clr, sizeof(ClearArgs));
        up(d_invy2 + s, &a.invy2, sizeof(StridedArgs));
        HIPCHK(hipMemcpyAsync(hsc + s, c->scal, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(ev_sc, st));
        HIPCHK(hipEventRecord(ev, st));
        slot[s] = c;
        rep = r ? r : c;
    }

    void detach(int s)
    {
        if (s < 0 || s >= nslots) throw Error(SPNS_EINVAL, "slot out of range");
        HIPCHK(hipStreamSynchronize(st));
        slot[s] = nullptr;
        bool any = false;
        for (spns_ctx* o : slot) any = any || o;
        if (!any) rep = nullptr;
        else if (rep == nullptr || std::find(slot.begin(), slot.end(), rep) == slot.end())
            for (spns_ctx* o : slot) if (o) { rep = o; break; }
    }

    // one step of every live slot: the single-run step's launches
    // (spns_ctx::enqueue_step), each once over all slots
    void enqueue_step()
    {
        spns_ctx* c = rep;
        const int p = c->prec;
        auto bl = [&](const void* a, const void* a2 = nullptr) { return BatchLaunch{a, a2, d_left, nslots}; };
        BatchLaunch b;
        b = bl(d_clr);
        HIPCHK(launch_clear(nullptr, nullptr, st, &b));
        b = bl(d_invy);
        HIPCHK(launch_strided(c->My, p, 1, p == PREC_FP64, +1, hs[slot_of(c)].invy, c->nk1, st, &b));
        b = bl(d_invx);
        HIPCHK(launch_strided(c->Mx, p, p == PREC_FP64, p == PREC_FP64, +1, hs[slot_of(c)].invx, c->My, st, &b));
        b = bl(d_z);
        HIPCHK(launch_zpass(c->Mz, p, 0, hs[slot_of(c)].z, c->grid_z, st, &b));
        b = bl(d_fwdx);
        HIPCHK(launch_strided(c->Mx, p, p == PREC_FP64, p == PREC_FP64, -1, hs[slot_of(c)].fwdx, c->My, st, &b));
        if (c->yinv) {
            b = bl(d_fwdy, d_epi);
            b.args3 = d_invy2;
            HIPCHK(launch_yepi(c->My, p, hs[slot_of(c)].fwdy, c->nk1, hs[slot_of(c)].epi, c->grid_yinv, st, &b,
                               &hs[slot_of(c)].invy2));
        } else if (c->grid_yepi > 0) {
            b = bl(d_fwdy, d_epi);
            HIPCHK(launch_yepi(c->My, p, hs[slot_of(c)].fwdy, c->nk1, hs[slot_of(c)].epi, c->grid_yepi, st, &b));
        } else {
            b = bl(d_fwdy);
            HIPCHK(launch_strided(c->My, p, p == PREC_FP64, p == PREC_FP64, -1, hs[slot_of(c)].fwdy, c->nk1, st, &b));
            b = bl(d_epi);
            HIPCHK(launch_epilogue(p, hs[slot_of(c)].epi, c->grid_epi, st, &b));
        }
        b = bl(d_fin);
        HIPCHK(launch_finalize(hs[slot_of(c)].fin, st, &b));
    }

    int slot_of(const spns_ctx* c) const
    {
        for (int s = 0; s < nslots; s++)
            if (slot[s] == c) return s;
        return 0;
    }

    void step(const int32_t* nsteps)
    {
        if (!rep) throw Error(SPNS_EINVAL, "no context attached");
        SlotSteps n{};
        for (int s = 0; s < nslots; s++) {
            const int v = slot[s] ? nsteps[s] : 0;
            if (v < 0 || v > chunk) throw Error(SPNS_EINVAL, "steps per slot must be in 0..chunk");
            n.n[s] = v;
        }
        HIPCHK(hipSetDevice(device));
        if (!graph) {
            // the chunk's launches read only the device arrays and counters,
            // so one capture serves every later set of slots
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (int i = 0; i < chunk; i++) enqueue_step();
            } catch (...) {
                hipGraph_t dummy;
                hipStreamEndCapture(st, &dummy);
                if (dummy) hipGraphDestroy(dummy);
                throw;
            }
            HIPCHK(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            HIPCHK(e);
        }
        HIPCHK(launch_set_left(d_left, n, nslots, st));
        HIPCHK(hipGraphLaunch(graph, st));
        for (int s = 0; s < nslots; s++)
            if (slot[s] && n.n[s] > 0) HIPCHK(hipMemcpyAsync(hsc + s, slot[s]->scal, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(ev_sc, st));
        HIPCHK(hipEventRecord(ev, st));
    }

    // the slots' scalar copies (hsc) have landed; fetches queued after them
    // may still be in flight
    bool scalars_ready()
    {
        const hipError_t e = hipEventQuery(ev_sc);
        if (e == hipErrorNotReady) {
            (void)hipGetLastError();
            return false;
        }
        HIPCHK(e);
        return true;
    }

    // the slot's state (compact layout) and spectrum copied into host
    // buffers on the batch's stream (asynchronous when they are pinned)
    void fetch(int s, double* state, double* ek)
    {
        if (s < 0 || s >= nslots || !slot[s]) throw Error(SPNS_EINVAL, "no context in this slot");
        spns_ctx* c = slot[s];
        if (state)
            for (int q = 0; q < 3; q++)
                HIPCHK(hipMemcpy2DAsync(state + 2 * (size_t)q * c->nmodes, sizeof(cx<double>) * c->nk3, c->uk[q],
                                        sizeof(cx<double>) * c->nk3p, sizeof(cx<double>) * c->nk3, (size_t)c->nk2 * c->nk1,
                                        hipMemcpyDeviceToHost, st));
        if (ek) HIPCHK(hipMemcpyAsync(ek, c->ek, sizeof(double) * c->nbins, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(ev, st));
    }
};

namespace {

int finish(spns_ctx* c, const std::exception& e, int code)
{
    if (c) c->err = e.what();
    return code;
}

thread_local std::string g_create_error;

}  // namespace

#define API_BEGIN(c) \
    if (!(c)) return SPNS_EINVAL; \
    try {
#define API_END(c)                                   \
    }                                                \
    catch (const Error& e) { return finish(c, e, e.code); } \
    catch (const std::exception& e) { return finish(c, e, SPNS_EINVAL); } \
    return SPNS_OK;

extern "C" {

int spns_create(const spns_config* cfg, spns_ctx** out)
{
    if (!cfg || !out) return SPNS_EINVAL;
    *out = nullptr;
    spns_ctx* c = new spns_ctx();
    try {
        c->init(*cfg);
        c->scratch_dev = noalloc_scratch(c->cfg.device, &c->scratch_err);  // device is current after init
    } catch (const Error& e) {
        g_create_error = e.what();
        c->release();
        delete c;
        return e.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        c->release();
        delete c;
        return SPNS_EINVAL;
    }
    *out = c;
    return SPNS_OK;
}

int spns_reconfigure(spns_ctx* ctx, const spns_config* cfg)
{
    API_BEGIN(ctx)
    if (!cfg) throw Error(SPNS_EINVAL, "null config");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    ctx->x1_stale();  // the state changes
    ctx->reconfigure(*cfg);
    API_END(ctx)
}

int spns_destroy(spns_ctx* ctx)
{
    if (!ctx) return SPNS_EINVAL;
    hipSetDevice(ctx->cfg.device);
    hipStreamSynchronize(ctx->st);
    ctx->release();
    delete ctx;
    return SPNS_OK;
}

int64_t spns_device_bytes(const spns_config* c)
{
    if (!c || c->Nx < 2 || c->Ny < 2 || c->Nz < 2) return -1;
    const int64_t Mx = c->Mx > 0 ? c->Mx : (3 * c->Nx) / 2, My = c->My > 0 ? c->My : (3 * c->Ny) / 2;
    const int64_t nk1 = 2 * (c->Nx / 2) + 1, nk2 = 2 * (c->Ny / 2) + 1, nk3 = c->Nz / 2 + 1;
    const int64_t nk3p = (nk3 + 15) / 16 * 16;
    const int64_t tsz = c->precision == SPNS_PREC_FP32_SPLIT ? 8 : 16;
    const int64_t nfwd = c->C_Smag > 0 ? 9 : 3;
    // state (and the step-start copy) in FP64, R history in the transform type
    const int64_t state = 16 * nk2 * nk1 * nk3p * ((c->isSelfAdaptive || c->gradient_invariants || c->keep_step_start) ? 6 : 3) +
                          tsz * nk2 * nk1 * nk3p * 3;
    const int64_t x1 = std::max<int64_t>(6, nfwd) * My * nk1 * nk3p * tsz;
    const int64_t x2 = 9 * My * Mx * nk3p * tsz;
    return state + x1 + x2 + (64 << 20);  // + partial sums, twiddles, scalars
}

int64_t spns_nmodes(const spns_ctx* ctx) { return ctx ? ctx->nmodes : -1; }
int32_t spns_nbins(const spns_ctx* ctx) { return ctx ? ctx->nbins : -1; }
int32_t spns_last_rad(const spns_ctx* ctx) { return ctx ? ctx->last_rad : -1; }
const char* spns_last_error(const spns_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int spns_set_state(spns_ctx* ctx, const double* state)
{
    API_BEGIN(ctx)
    ctx->x1_stale();  // the state changes
    if (!state) throw Error(SPNS_EINVAL, "null state");
    std::vector<std::complex<double>> h(3 * (size_t)ctx->nmodes);
    std::memcpy(h.data(), state, sizeof(double) * 6 * ctx->nmodes);
    ctx->upload_state(h);
    API_END(ctx)
}

int spns_get_state(spns_ctx* ctx, double* state)
{
    API_BEGIN(ctx)
    if (!state) throw Error(SPNS_EINVAL, "null output");
    auto h = ctx->download_state(ctx->uk);
    std::memcpy(state, h.data(), sizeof(double) * 6 * ctx->nmodes);
    API_END(ctx)
}

int spns_get_state_planes(spns_ctx* ctx, int32_t i2_first, int32_t n_i2, double* out)
{
    API_BEGIN(ctx)
    if (!out || i2_first < 0 || n_i2 < 0 || (int64_t)i2_first + n_i2 > ctx->nk2)
        throw Error(SPNS_EINVAL, "i2 range outside the compact box");
    const size_t plane = (size_t)ctx->nk1 * ctx->nk3;  // modes per i2 plane (compact)
    for (int c = 0; c < 3; c++)
        HIPCHK(hipMemcpy2DAsync(out + 2 * (size_t)c * n_i2 * plane, sizeof(cx<double>) * ctx->nk3,
                                ctx->uk[c] + (size_t)i2_first * ctx->nk1 * ctx->nk3p, sizeof(cx<double>) * ctx->nk3p,
                                sizeof(cx<double>) * ctx->nk3, (size_t)n_i2 * ctx->nk1, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    API_END(ctx)
}

int spns_get_rhist(spns_ctx* ctx, double* rh)
{
    API_BEGIN(ctx)
    if (!rh) throw Error(SPNS_EINVAL, "null output");
    if (ctx->prec == PREC_FP64) {
        auto h = ctx->download_state((cx<double>* const*)ctx->R);
        std::memcpy(rh, h.data(), sizeof(double) * 6 * ctx->nmodes);
    } else {
        // FP32-split: the history is held in FP32; widened here
        std::vector<cx<float>> h(3 * (size_t)ctx->nmodes);
        for (int c = 0; c < 3; c++)
            HIPCHK(hipMemcpy2DAsync(h.data() + c * ctx->nmodes, sizeof(cx<float>) * ctx->nk3, ctx->R[c],
                                    sizeof(cx<float>) * ctx->nk3p, sizeof(cx<float>) * ctx->nk3, (size_t)ctx->nk2 * ctx->nk1,
                                    hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        for (size_t i = 0; i < h.size(); i++) {
            rh[2 * i] = h[i].x;
            rh[2 * i + 1] = h[i].y;
        }
    }
    API_END(ctx)
}

int spns_init_from_physical(spns_ctx* ctx, const double* u, const double* v, const double* w, int32_t nxf, int32_t nyf,
                            int32_t nzf, double* ek_file, double* ens_file)
{
    API_BEGIN(ctx)
    ctx->x1_stale();  // the state changes
    if (!u || !v || !w || nxf < 2 || nyf < 2 || nzf < 2) throw Error(SPNS_EINVAL, "bad file field");
    // r2c of the file grid on the device (src/HIT_LES.cpp:557-583)
    spns_fft* f = nullptr;
    int rc = spns_fft_create(nxf, nyf, nzf, SPNS_PREC_FP64, ctx->cfg.device, &f);
    if (rc) throw Error(rc, "file grid: FFT length not compiled in");
    const int nzh = nzf / 2 + 1;
    const size_t nr = (size_t)nxf * nyf * 2 * nzh, ncf = (size_t)nyf * nxf * nzh;
    std::vector<double> pad(nr, 0.0);
    std::vector<std::complex<double>> spec[3];
    double *dr = nullptr, *dc = nullptr;
    try {
        HIPCHK(hipMalloc((void**)&dr, sizeof(double) * nr));
        HIPCHK(hipMalloc((void**)&dc, sizeof(double) * 2 * ncf));
        const double* src[3] = {u, v, w};
        for (int c = 0; c < 3; c++) {
            for (int i = 0; i < nxf; i++)
                for (int j = 0; j < nyf; j++)
                    for (int k = 0; k < nzf; k++)
                        pad[((size_t)i * nyf + j) * 2 * nzh + k] = src[c][((size_t)i * nyf + j) * nzf + k];
            HIPCHK(hipMemcpy(dr, pad.data(), sizeof(double) * nr, hipMemcpyHostToDevice));
            if ((rc = spns_fft_r2c(f, dr, dc))) throw Error(rc, spns_fft_last_error(f));
            spec[c].resize(ncf);
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipMemcpy(spec[c].data(), dc, sizeof(double) * 2 * ncf, hipMemcpyDeviceToHost));
            const double norm = (double)(nxf * nyf * nzf);
            for (auto& z : spec[c]) z = std::complex<double>(z.real() / norm, z.imag() / norm);
        }
    } catch (...) {
        hipFree(dr); hipFree(dc);
        spns_fft_destroy(f);
        throw;
    }
    hipFree(dr); hipFree(dc);
    spns_fft_destroy(f);
    // file spectrum energy (src/HIT_LES.cpp:1268-1294); spectrum layout [y][x][z]
    double ekf = 0.0, ensf = 0.0;
    for (int a = 0; a < nxf; a++)
        for (int b = 0; b < nyf; b++)
            for (int k3 = 0; k3 < nzh; k3++) {
                size_t id = ((size_t)b * nxf + a) * nzh + k3;
                int k1 = ffreq(a, nxf), k2 = ffreq(b, nyf);
                int r2 = k1 * k1 + k2 * k2 + k3 * k3;
                if (!((k1 % ctx->cfg.Lx == 0) && (k2 % ctx->cfg.Ly == 0) && (k3 % ctx->cfg.Lz == 0))) continue;
                bool conj = (k3 != 0) && !((nzf % 2 == 0) && (k3 == nzf / 2));
                double e = (spec[0][id].real() * spec[0][id].real()) + (spec[0][id].imag() * spec[0][id].imag());
                e += (spec[1][id].real() * spec[1][id].real()) + (spec[1][id].imag() * spec[1][id].imag());
                e += (spec[2][id].real() * spec[2][id].real()) + (spec[2][id].imag() * spec[2][id].imag());
                e *= 0.5;
                if (conj) { ekf += 2 * e; ensf += 2 * (2 * r2 * e); }
                else { ekf += e; ensf += (2 * r2 * e); }
            }
    if (ek_file) *ek_file = ekf;
    if (ens_file) *ens_file = ensf;
    // truncation (src/HIT_LES.cpp:738-763)
    std::vector<std::complex<double>> h(3 * (size_t)ctx->nmodes, 0.0);
    for (int k2 = -ctx->Ny_2; k2 <= ctx->Ny_2; k2++)
        for (int k1 = -ctx->Nx_2; k1 <= ctx->Nx_2; k1++)
            for (int k3 = 0; k3 <= ctx->Nz_2; k3++) {
                if (!ctx->retained(k1, k2, k3)) continue;
                if (!((std::abs(k1) <= (nxf - 1) / 2) && (std::abs(k2) <= (nyf - 1) / 2) && (k3 <= (nzf - 1) / 2))) continue;
                int af = k1 >= 0 ? k1 : nxf + k1, bf = k2 >= 0 ? k2 : nyf + k2;
                size_t idf = ((size_t)bf * nxf + af) * nzh + k3;
                int64_t q = ctx->compact_index(k1, k2, k3);
                for (int c = 0; c < 3; c++) h[c * ctx->nmodes + q] = spec[c][idf];
            }
    ctx->upload_state(h);
    API_END(ctx)
}

int spns_init_k41(spns_ctx* ctx, int32_t nranks)
{
    API_BEGIN(ctx)
    ctx->x1_stale();  // the state changes
    if (nranks < 1) throw Error(SPNS_EINVAL, "nranks must be >= 1");
    const int Mx = ctx->Mx, My = ctx->My, Mz2 = ctx->Mz / 2;
    std::vector<std::complex<double>> h(3 * (size_t)ctx->nmodes, 0.0);
    const int block = (My + nranks - 1) / nranks;
    // every rank draws over its own y-planes (src/HIT_LES.cpp:504-525)
    for (int r = 0; r < nranks; r++) {
        const int b0 = r * block, b1 = std::min(My, b0 + block);
        srand(0);
        for (int b = b0; b < b1; b++)
            for (int a = 0; a < Mx; a++)
                for (int k3 = 0; k3 <= Mz2; k3++) {
                    const int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
                    const bool d = ctx->retained(k1, k2, k3);
                    double val[2][3] = {{0, 0, 0}, {0, 0, 0}};
                    for (int ic = 0; ic <= 1; ic++) {
                        if (d) {
                            int r2 = k1 * k1 + k2 * k2 + k3 * k3;
                            int K = (int)std::lround(std::sqrt((double)r2));
                            if (K <= ctx->last_rad)
                                for (int c = 0; c < 3; c++) val[ic][c] = (((double)rand() / RAND_MAX) * 2.0 - 1.0);
                        }
                    }
                    if (d) {
                        int64_t q = ctx->compact_index(k1, k2, k3);
                        for (int c = 0; c < 3; c++) h[c * ctx->nmodes + q] = std::complex<double>(val[0][c], val[1][c]);
                    }
                }
    }
    ctx->upload_state(h);
    HIPCHK(launch_ccc(ctx->shell_args(), ctx->st));  // src/HIT_LES.cpp:527
    std::vector<double> k41(ctx->last_rad + 1, 0.0);
    for (int K = 1; K <= ctx->last_rad; K++) k41[K] = std::pow(K, -5.0 / 3.0);
    ctx->forcing_energy_cascade(k41.data(), ctx->last_rad, 0);
    API_END(ctx)
}

int spns_init_random(spns_ctx* ctx, uint64_t seed, double k0, double urms)
{
    API_BEGIN(ctx)
    ctx->x1_stale();  // the state changes
    if (!(k0 > 0) || !(urms >= 0)) throw Error(SPNS_EINVAL, "k0 must be > 0 and urms >= 0");
    RandArgs r{};
    for (int c = 0; c < 3; c++) r.uk[c] = ctx->uk[c];
    r.nk1 = ctx->nk1; r.nk2 = ctx->nk2; r.nk3 = ctx->nk3; r.nk3p = ctx->nk3p; r.nmodes = ctx->nmodes_p;
    r.notcubic = ctx->cfg.isNotCubic; r.Lx = ctx->cfg.Lx; r.Ly = ctx->cfg.Ly; r.Lz = ctx->cfg.Lz;
    r.seed = seed;
    r.k0 = k0;
    HIPCHK(launch_init_random(r, ctx->st));
    ctx->reduce_state();
    DevScalars s = ctx->read_scalars();
    const double target = 1.5 * urms * urms;  // u_rms^2 = <u_i u_i>/3
    const double sc = s.ek > 0 ? std::sqrt(target / s.ek) : 0.0;
    HIPCHK(launch_scale_state(ctx->shell_args(), sc, ctx->st));
    ctx->reset_history();
    ctx->reduce_state();
    API_END(ctx)
}

int spns_finish_init(spns_ctx* ctx, int32_t isReLambda)
{
    API_BEGIN(ctx)
    ctx->reset_history();
    DevScalars s = ctx->read_scalars();
    s.nu = ctx->cfg.nu;
    s.kappa = 0.5;
    ctx->write_scalars(s);
    ctx->velocity_pass(nullptr);
    if (!isReLambda) HIPCHK(launch_finalize(ctx->fin_args(FIN_DT), ctx->st));
    ctx->reduce_state();
    HIPCHK(hipStreamSynchronize(ctx->st));
    API_END(ctx)
}

int spns_step(spns_ctx* ctx, int32_t nsteps)
{
    API_BEGIN(ctx)
    if (nsteps < 0) throw Error(SPNS_EINVAL, "nsteps < 0");
    ctx->step(nsteps);
    API_END(ctx)
}

int spns_recalculate_timestep(spns_ctx* ctx)
{
    API_BEGIN(ctx)
    HIPCHK(launch_finalize(ctx->fin_args(FIN_DT), ctx->st));
    API_END(ctx)
}

int spns_set_nu(spns_ctx* ctx, double nu)
{
    API_BEGIN(ctx)
    DevScalars s = ctx->read_scalars();
    s.nu = nu;
    ctx->cfg.nu = nu;
    ctx->write_scalars(s);
    API_END(ctx)
}

int spns_forcing_reynolds_lambda(spns_ctx* ctx, double relambda)
{
    API_BEGIN(ctx)
    DevScalars s = ctx->read_scalars();
    s.nu = s.ek / relambda * std::sqrt(20.0 / 3.0 / s.om);  // src/HIT_LES.cpp:488
    ctx->cfg.nu = s.nu;
    ctx->write_scalars(s);
    API_END(ctx)
}

int spns_recalculate_reynolds_lambda(spns_ctx* ctx, double* relambda)
{
    API_BEGIN(ctx)
    DevScalars s = ctx->read_scalars();
    double r = s.ek / s.nu * std::sqrt(20.0 / 3.0 / s.om);
    if (relambda) *relambda = r;
    API_END(ctx)
}

int spns_forcing_energy_cascade(spns_ctx* ctx, const double* fek, int64_t last_input_rad, int32_t nullify)
{
    API_BEGIN(ctx)
    ctx->x1_stale();  // the state changes
    if (!fek || last_input_rad < 0) throw Error(SPNS_EINVAL, "bad forced spectrum");
    ctx->forcing_energy_cascade(fek, last_input_rad, nullify);
    API_END(ctx)
}

int spns_complex_conjugate_correction(spns_ctx* ctx)
{
    API_BEGIN(ctx)
    ctx->x1_stale();  // the state changes
    HIPCHK(launch_ccc(ctx->shell_args(), ctx->st));
    ctx->reduce_state();
    API_END(ctx)
}

int spns_energy_cascade(spns_ctx* ctx, double* ekh, int32_t nbins, double* ek_tot)
{
    API_BEGIN(ctx)
    if (ekh && nbins > ctx->nbins) throw Error(SPNS_EINVAL, "nbins exceeds the spectrum");
    if (ekh && nbins > 0) {
        HIPCHK(hipMemcpyAsync(ekh, ctx->ek, sizeof(double) * nbins, hipMemcpyDeviceToHost, ctx->st));
    }
    DevScalars s = ctx->read_scalars();
    if (ek_tot) *ek_tot = s.ek;
    API_END(ctx)
}

// spns_scalars from the device scalars; SPNS_ENONFINITE if the run diverged
static int zpass_kernel_of(int L, int prec, const ZArgs& z) { return zpass_kernel(L, prec, z); }

static void to_scalars(const spns_ctx* ctx, const DevScalars& s, spns_scalars* out)
{
    out->time = s.time;
    out->dt = s.dt;
    out->nu = s.nu;
    out->maxvel = ctx->sa ? s.maxvel_sa : s.maxvel;
    out->ek = s.ek;
    out->om = s.om;
    out->relambda = s.ek / s.nu * std::sqrt(20.0 / 3.0 / s.om);
    const double P = ctx->nf;
    const double m2 = s.skew2 / P, m3 = s.skew3 / P;
    out->skewness = m2 > 0 ? m3 / std::pow(m2, 1.5) : 0.0;
    out->kappa = s.kappa;
    out->iter = s.iter;
    if (!std::isfinite(s.maxvel) || !std::isfinite(s.ek) || !std::isfinite(s.om) || !std::isfinite(s.dt))
        throw Error(SPNS_ENONFINITE, "non-finite velocity field after step " + std::to_string(s.iter));
}

int spns_pass_kernels(spns_ctx* ctx, int32_t* zpass_kernel, int32_t* fused_epilogue)
{
    API_BEGIN(ctx)
    if (zpass_kernel) {
        ZArgs z = ctx->zpass_args(nullptr);
        z.kernel = ctx->zkernel;
        *zpass_kernel = zpass_kernel_of(ctx->Mz, ctx->prec, z);
    }
    if (fused_epilogue) *fused_epilogue = ctx->yinv ? 2 : (ctx->grid_yepi > 0 ? 1 : 0);
    API_END(ctx)
}

int spns_get_scalars(spns_ctx* ctx, spns_scalars* out)
{
    API_BEGIN(ctx)
    if (!out) throw Error(SPNS_EINVAL, "null output");
    to_scalars(ctx, ctx->read_scalars(), out);
    API_END(ctx)
}

// The velocity pass (current state) or the step-start pass into a device
// buffer, then [c][y][x][z] copied out as [c][x][y][z] doubles.
static void physical_out(spns_ctx* ctx, double* uvw, bool step_start)
{
    if (!uvw) throw Error(SPNS_EINVAL, "null output");
    const int Mx = ctx->Mx, My = ctx->My, Mz = ctx->Mz;
    const size_t P = (size_t)Mx * My * Mz;
    const size_t ts = ctx->prec == PREC_FP64 ? 8 : 4;
    void* d = nullptr;
    ctx->alloc(&d, 3 * P * ts);
    try {
        if (step_start)
            ctx->velocity_step_start(d);
        else
            ctx->velocity_pass(d);
        std::vector<char> hbuf(3 * P * ts);
        HIPCHK(hipMemcpyAsync(hbuf.data(), d, 3 * P * ts, hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        for (int c = 0; c < 3; c++)
            for (int y = 0; y < My; y++)
                for (int x = 0; x < Mx; x++)
                    for (int z = 0; z < Mz; z++) {
                        size_t s = c * P + ((size_t)y * Mx + x) * Mz + z;
                        double v = ts == 8 ? ((double*)hbuf.data())[s] : (double)((float*)hbuf.data())[s];
                        uvw[c * P + ((size_t)x * My + y) * Mz + z] = v;
                    }
    } catch (...) {
        hipFree(d);
        throw;
    }
    hipFree(d);
}

int spns_velocity_physical(spns_ctx* ctx, double* uvw)
{
    API_BEGIN(ctx)
    physical_out(ctx, uvw, false);
    API_END(ctx)
}

int spns_velocity_physical_step_start(spns_ctx* ctx, double* uvw)
{
    API_BEGIN(ctx)
    physical_out(ctx, uvw, true);
    API_END(ctx)
}

int spns_check_phys_four_energy(spns_ctx* ctx, double atol, double* ek_real, double* ek_four, double* ek_spns)
{
    API_BEGIN(ctx)
    double er = 0.0, ef = 0.0, es = 0.0;
    ctx->check_phys_four_energy(&er, &ef, &es);
    if (ek_real) *ek_real = er;
    if (ek_four) *ek_four = ef;
    if (ek_spns) *ek_spns = es;
    const double d1 = std::fabs(er - ef), d2 = std::fabs(es - ef);
    if (!(d1 <= atol) || !(d2 <= atol)) {  // src/HIT_LES.cpp:1246-1250
        char msg[256];
        std::snprintf(msg, sizeof msg, "Ek differences exceed tolerance [EkReal vs EkFour %e, EkSpNS vs EkFour: %e]", d1, d2);
        throw Error(SPNS_ECHECK, msg);
    }
    API_END(ctx)
}

int spns_recalculate_velocity_antitransform(spns_ctx* ctx)
{
    API_BEGIN(ctx)
    ctx->velocity_pass(nullptr);
    API_END(ctx)
}

int spns_synchronize(spns_ctx* ctx)
{
    API_BEGIN(ctx)
    HIPCHK(hipStreamSynchronize(ctx->st));
    API_END(ctx)
}

int spns_query(spns_ctx* ctx, int32_t* done)
{
    API_BEGIN(ctx)
    if (!done) throw Error(SPNS_EINVAL, "done is null");
    const hipError_t e = hipEventQuery(ctx->ev1);
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: clear the sticky status
        *done = 0;
    } else {
        HIPCHK(e);
        *done = 1;
        // the stream is idle: a small read tells whether the run diverged
        const DevScalars s = ctx->read_scalars();
        if (!std::isfinite(s.maxvel) || !std::isfinite(s.ek) || !std::isfinite(s.om) || !std::isfinite(s.dt))
            throw Error(SPNS_ENONFINITE, "non-finite velocity field after step " + std::to_string(s.iter));
    }
    API_END(ctx)
}

int spns_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gb_per_s)
{
    if (!gb_per_s || bytes <= 0 || bytes % 16 || reps < 1) return SPNS_EINVAL;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPNS_OK;
    try {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipMalloc(&a, bytes));
        HIPCHK(hipMalloc(&b, bytes));
        HIPCHK(hipMemsetAsync(a, 0, bytes, st));
        int ncu = 0;
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        const int blocks = ncu * 16;
        HIPCHK(launch_stream_copy(a, b, bytes, blocks, st));  // warm-up
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) HIPCHK(launch_stream_copy((r & 1) ? b : a, (r & 1) ? a : b, bytes, blocks, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *gb_per_s = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;  // bytes read + written
    } catch (const Error& e) {
        g_create_error = e.what();
        rc = e.code;
    }
    if (st) hipStreamSynchronize(st);
    hipFree(a); hipFree(b);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    return rc;
}

int spns_device_memory(int32_t device, int64_t* free_bytes, int64_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return SPNS_EINVAL;
    size_t fr = 0, tot = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        g_create_error = "hipMemGetInfo failed on device " + std::to_string(device);
        return SPNS_EHIP;
    }
    *free_bytes = (int64_t)fr;
    *total_bytes = (int64_t)tot;
    return SPNS_OK;
}

int spns_invariant_spectra(spns_ctx* ctx, int32_t nbins, double* spectra)
{
    API_BEGIN(ctx)
    if (!spectra || nbins < ctx->nbins) throw Error(SPNS_EINVAL, "spectra must hold 5 x spns_nbins values");
    std::vector<double> h(5 * (size_t)ctx->nbins);
    ctx->invariant_spectra(h.data());
    for (int f = 0; f < 5; f++)
        for (int K = 0; K < ctx->nbins; K++) spectra[(size_t)f * nbins + K] = h[(size_t)f * ctx->nbins + K];
    API_END(ctx)
}

int spns_invariant_spectra_noalloc(spns_ctx* ctx, int32_t nbins, double* spectra)
{
    API_BEGIN(ctx)
    if (!spectra || nbins < ctx->nbins) throw Error(SPNS_EINVAL, "spectra must hold 5 x spns_nbins values");
    const size_t need = ctx->inv_spectra_scratch();
    if (need > kNoAllocScratchDoubles)
        throw Error(SPNS_EINVAL, "invariant spectra of this grid need " + std::to_string(need) +
                                     " doubles of scratch; the fixed scratch holds " +
                                     std::to_string(kNoAllocScratchDoubles) + " (use spns_invariant_spectra)");
    if (!ctx->scratch_dev)
        throw Error(SPNS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(ctx->scratch_err));
    HIPCHK(hipSetDevice(ctx->cfg.device));
    ctx->invariant_spectra_dev(ctx->scratch_dev);
    HIPCHK(hipMemcpy2DAsync(spectra, sizeof(double) * nbins, ctx->scratch_dev, sizeof(double) * ctx->nbins,
                            sizeof(double) * ctx->nbins, 5, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    API_END(ctx)
}

int spns_profile_passes(spns_ctx* ctx, int32_t nsteps, double* ms)
{
    API_BEGIN(ctx)
    if (nsteps < 1 || !ms) throw Error(SPNS_EINVAL, "nsteps >= 1 and an output array are required");
    constexpr int NP = 8;
    hipEvent_t ev[NP + 1];
    for (int i = 0; i <= NP; i++) HIPCHK(hipEventCreate(&ev[i]));
    double acc[NP] = {0};
    try {
        for (int s = 0; s < nsteps; s++) {
            HIPCHK(hipEventRecord(ev[0], ctx->st));
            HIPCHK(launch_clear(ctx->mv, ctx->mv + 1, ctx->st));
            HIPCHK(hipEventRecord(ev[1], ctx->st));
            ctx->inv_y(false);
            HIPCHK(hipEventRecord(ev[2], ctx->st));
            ctx->inv_x(false);
            HIPCHK(hipEventRecord(ev[3], ctx->st));
            ctx->zpass(0, nullptr);
            HIPCHK(hipEventRecord(ev[4], ctx->st));
            ctx->fwd_x();
            HIPCHK(hipEventRecord(ev[5], ctx->st));
            FinArgs f = ctx->fin_args(FIN_REDUCE | FIN_SKEW | FIN_MAXVEL | FIN_DT | FIN_ADVANCE);
            if (ctx->grid_yepi > 0) {
                // fused: the whole kernel is timed as fwd_y, the epilogue slot stays 0
                f.nblk_epi = ctx->fwd_y_epilogue();
                HIPCHK(hipEventRecord(ev[6], ctx->st));
            } else {
                ctx->fwd_y();
                HIPCHK(hipEventRecord(ev[6], ctx->st));
                HIPCHK(launch_epilogue(ctx->prec, ctx->epi_args(EPI_STEP), ctx->grid_epi, ctx->st));
            }
            HIPCHK(hipEventRecord(ev[7], ctx->st));
            HIPCHK(launch_finalize(f, ctx->st));
            HIPCHK(hipEventRecord(ev[8], ctx->st));
            HIPCHK(hipEventSynchronize(ev[8]));
            for (int i = 0; i < NP; i++) {
                float f = 0.f;
                HIPCHK(hipEventElapsedTime(&f, ev[i], ev[i + 1]));
                acc[i] += f;
            }
        }
    } catch (...) {
        for (int i = 0; i <= NP; i++) hipEventDestroy(ev[i]);
        throw;
    }
    for (int i = 0; i <= NP; i++) hipEventDestroy(ev[i]);
    for (int i = 0; i < NP; i++) ms[i] = acc[i] / nsteps;
    API_END(ctx)
}

int spns_last_step_ms(spns_ctx* ctx, double* ms)
{
    API_BEGIN(ctx)
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    if (ms) *ms = f;
    API_END(ctx)
}

// ------------------------------ FFT_package ------------------------------
int spns_fft_create(int32_t Mx, int32_t My, int32_t Mz, int32_t precision, int32_t device, spns_fft** out)
{
    if (!out) return SPNS_EINVAL;
    *out = nullptr;
    if (!supported_length(Mx) || !supported_length(My) || !supported_length(Mz)) {
        g_create_error = "FFT length not compiled in";
        return SPNS_EUNSUPPORTED;
    }
    spns_fft* f = new spns_fft();
    try {
        f->Mx = Mx; f->My = My; f->Mz = Mz; f->device = device;
        f->prec = precision == SPNS_PREC_FP32_SPLIT ? PREC_FP32 : PREC_FP64;
        HIPCHK(hipSetDevice(device));
        // a blocking stream: work the caller queued on the legacy default
        // stream (e.g. the fill of an input) is ordered before each transform
        HIPCHK(hipStreamCreateWithFlags(&f->st, hipStreamDefault));
        f->twx = upload_twiddles(Mx, f->prec);
        f->twy = upload_twiddles(My, f->prec);
        f->twz = upload_twiddles(Mz, f->prec);
    } catch (const Error& e) {
        g_create_error = e.what();
        spns_fft_destroy(f);
        return e.code;
    }
    *out = f;
    return SPNS_OK;
}

int spns_fft_destroy(spns_fft* f)
{
    if (!f) return SPNS_EINVAL;
    if (f->st) hipStreamSynchronize(f->st);
    hipFree(f->twx); hipFree(f->twy); hipFree(f->twz);
    if (f->st) hipStreamDestroy(f->st);
    delete f;
    return SPNS_OK;
}

// The transforms return with their output complete, like the FFTW plans
// they replace (src/FFTW_package.cpp:7-14 call fftw_execute, which returns
// when done).
int spns_fft_r2c(spns_fft* f, void* rin, void* cout)
{
    if (!f || !rin || !cout) return SPNS_EINVAL;
    try {
        HIPCHK(hipSetDevice(f->device));
        f->z(rin, cout, -1);
        f->fft_y(cout, -1);
        f->fft_x(cout, -1);
        HIPCHK(hipStreamSynchronize(f->st));
    } catch (const Error& e) {
        f->err = e.what();
        return e.code;
    }
    return SPNS_OK;
}

int spns_fft_c2r(spns_fft* f, void* cin, void* rout)
{
    if (!f || !cin || !rout) return SPNS_EINVAL;
    try {
        HIPCHK(hipSetDevice(f->device));
        f->fft_x(cin, +1);
        f->fft_y(cin, +1);
        f->z(cin, rout, +1);
        HIPCHK(hipStreamSynchronize(f->st));
    } catch (const Error& e) {
        f->err = e.what();
        return e.code;
    }
    return SPNS_OK;
}

int spns_fft_normalize(spns_fft* f, void* c, int32_t Nx, int32_t Ny, int32_t Nz)
{
    if (!f || !c) return SPNS_EINVAL;
    try {
        NormArgs a{};
        a.data = c;
        a.Mx = f->Mx; a.My = f->My; a.Mz = f->Mz;
        a.Nx2 = Nx / 2; a.Ny2 = Ny / 2; a.Nz2 = Nz / 2;
        a.all = Nx <= 0;
        a.Lx = a.Ly = a.Lz = 1;
        a.nf = (double)((int64_t)f->Mx * f->My * f->Mz);
        HIPCHK(hipSetDevice(f->device));
        HIPCHK(launch_normalize(f->prec, a, f->st));
        HIPCHK(hipStreamSynchronize(f->st));
    } catch (const Error& e) {
        f->err = e.what();
        return e.code;
    }
    return SPNS_OK;
}

int spns_fft_synchronize(spns_fft* f)
{
    if (!f) return SPNS_EINVAL;
    hipError_t e = hipStreamSynchronize(f->st);
    if (e != hipSuccess) {
        f->err = hipGetErrorString(e);
        return SPNS_EHIP;
    }
    return SPNS_OK;
}

const char* spns_fft_last_error(const spns_fft* f) { return f ? f->err.c_str() : g_create_error.c_str(); }

// ------------------------------ batched stepping ------------------------------
#define BATCH_BEGIN(b) \
    if (!(b)) return SPNS_EINVAL; \
    try {
#define BATCH_END(b)                                                   \
    }                                                                  \
    catch (const Error& e) { (b)->err = e.what(); return e.code; }     \
    catch (const std::exception& e) { (b)->err = e.what(); return SPNS_EINVAL; } \
    return SPNS_OK;

int spns_batch_create(int32_t device, int32_t nslots, int32_t chunk, spns_batch** out)
{
    if (!out) return SPNS_EINVAL;
    *out = nullptr;
    spns_batch* b = new spns_batch();
    try {
        b->init(device, nslots, chunk);
    } catch (const Error& e) {
        g_create_error = e.what();
        b->release();
        delete b;
        return e.code;
    }
    *out = b;
    return SPNS_OK;
}

int spns_batch_destroy(spns_batch* b)
{
    if (!b) return SPNS_EINVAL;
    hipSetDevice(b->device);
    b->release();
    delete b;
    return SPNS_OK;
}

int spns_batch_attach(spns_batch* b, int32_t slot, spns_ctx* ctx)
{
    BATCH_BEGIN(b)
    HIPCHK(hipSetDevice(b->device));
    b->attach(slot, ctx);
    BATCH_END(b)
}

int spns_batch_detach(spns_batch* b, int32_t slot)
{
    BATCH_BEGIN(b)
    HIPCHK(hipSetDevice(b->device));
    b->detach(slot);
    BATCH_END(b)
}

int spns_batch_step(spns_batch* b, const int32_t* nsteps)
{
    BATCH_BEGIN(b)
    if (!nsteps) throw Error(SPNS_EINVAL, "nsteps is null");
    b->step(nsteps);
    BATCH_END(b)
}

int spns_batch_query(spns_batch* b, int32_t* done)
{
    BATCH_BEGIN(b)
    if (!done) throw Error(SPNS_EINVAL, "done is null");
    const hipError_t e = hipEventQuery(b->ev);
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
    } else {
        HIPCHK(e);
        *done = 1;
    }
    BATCH_END(b)
}

int spns_batch_synchronize(spns_batch* b)
{
    BATCH_BEGIN(b)
    HIPCHK(hipStreamSynchronize(b->st));
    BATCH_END(b)
}

int spns_batch_slot_scalars(spns_batch* b, int32_t slot, spns_scalars* out)
{
    BATCH_BEGIN(b)
    if (!out || slot < 0 || slot >= b->nslots || !b->slot[slot]) throw Error(SPNS_EINVAL, "no context in this slot");
    if (!b->scalars_ready()) throw Error(SPNS_EINVAL, "the batch is still running");
    to_scalars(b->slot[slot], b->hsc[slot], out);
    BATCH_END(b)
}

int spns_batch_fetch(spns_batch* b, int32_t slot, double* state, double* ek)
{
    BATCH_BEGIN(b)
    HIPCHK(hipSetDevice(b->device));
    b->fetch(slot, state, ek);
    BATCH_END(b)
}

const char* spns_batch_last_error(const spns_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int spns_host_alloc(int64_t bytes, void** out)
{
    if (!out || bytes <= 0) return SPNS_EINVAL;
    *out = nullptr;
    if (hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        g_create_error = "hipHostMalloc of " + std::to_string(bytes) + " bytes failed";
        return SPNS_ENOMEM;
    }
    return SPNS_OK;
}

int spns_host_free(void* p)
{
    if (!p) return SPNS_EINVAL;
    return hipHostFree(p) == hipSuccess ? SPNS_OK : SPNS_EHIP;
}

}  // extern "C"
