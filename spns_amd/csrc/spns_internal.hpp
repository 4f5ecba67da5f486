// Internal interfaces between the host context and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spns {

template <typename T>
struct cx {
    T x, y;
};

enum { PREC_FP64 = 0, PREC_FP32 = 1 };

// Scalars that live in device memory so a captured step needs no host sync.
struct DevScalars {
    double dt, time, nu, alpha, kappa;
    double maxvel, maxvel_sa;
    double ek, om, relambda;
    double skew2, skew3;
    int64_t iter;
    double ek_phys;  // sum of u.u over the FFT grid of the last velocity pass (FIN_EPHYS)
    double ek_four;  // E summed directly over the retained modes (FIN_EFOUR)
    double pad[1];
};

struct StridedJob {
    const void* in;
    void* out;
    int deriv;  // 1: multiply the input by i*k of the transformed dimension
    int pad;
    void* out2;  // if set: out gets the plain transform and out2 the i*k one, from one read of in
};

struct StridedArgs {
    StridedJob job[9];
    int njobs;
    int B;     // contiguous batch width (columns)
    int nin;   // stored input rows (odd, wrapped order) or L
    int nout;  // stored output rows (odd, wrapped order) or L
    int A;     // outer extent (set by the launcher)
    int padded_ok;  // column tiles rounded up to the tile width stay inside every input buffer
    int64_t in_a_stride, in_row_stride, out_a_stride, out_row_stride;  // in complex elements
    const void* tw;
    const int* skip;  // if set and *skip != 0 the launch does nothing (its outputs are already current)
    int* setflag;     // batched k_yepi_inv: as the inverse pass's arguments, the slot's flag the kernel sets
};

enum { ZK_WAVE = 0, ZK_BLOCK = 1, ZK_STAGED = 2, ZK_STAGED_NOFUSE = 3 };

struct ZArgs {
    void* f[9];  // X2 fields: u v w | ux vx wx | uy vy wy ; pencil offset pen*pitch
    int nk3, pitch;
    int les;
    int sa;  // also reduce the self-adaptive sum form of the CFL velocity
    int kernel;  // ZK_WAVE, ZK_BLOCK, ZK_STAGED (DNS fused where it fits) or ZK_STAGED_NOFUSE
    int64_t npencil;
    double Ax, Ay, Az;  // CFL divisors (src/HIT_LES.h:56-58)
    double ct;          // (C_Smag*delta)^2 (src/HIT_LES.cpp:132)
    unsigned long long* maxvel_bits;
    unsigned long long* maxvel_sa_bits;
    double* skew_partial;  // [gridDim.x][2]: MODE 0 sums of (du/dx)^2, (du/dx)^3; MODE 1 sum of u.u, 0
    void* phys;            // MODE 1: optional [3][npencil][L] physical export
    const void* tw;
};

struct ZPlainArgs {
    const void* in;
    void* out;
    int Mx, My;
    const void* tw;
};

enum { EPI_STEP = 0, EPI_REDUCE = 1 };

struct EpiArgs {
    const void* conv[9];  // 3 convective (+6 SGS for LES) transformed fields, compact box
    cx<double>* uk[3];
    void* R[3];  // R history: cx<T> of the kernel's transform type T (FP64 form: double; FP32-split: float)
    cx<double>* ukprev[3];
    int nk1, nk2, nk3, nk3p, nbins;
    int64_t nmodes;  // padded: nk2 * nk1 * nk3p
    int notcubic, Lx, Ly, Lz, Mz;
    int mode, les, rot, force_on, kfmin, kfmax, sa;
    int keepprev;  // store the start-of-step state in ukprev (without the kappa-scheme)
    double nf;  // Mx*My*Mz
    double ox, oy, oz;
    const DevScalars* scal;
    double* hist_partial;  // [grid][nbins]
    double* om_partial;    // [grid]
    double* ed_partial;    // [grid] E summed directly, EPI_REDUCE only (nullptr: not kept)
};

struct InvSpecArgs {
    const void* f[5];  // compact transformed invariants QG QS RG RS V2 (unnormalised)
    int nk1, nk2, nk3, nk3p, nbins;
    int notcubic, Lx, Ly, Lz, Mz;
    double inv_nf2;    // 1 / (Mx My Mz)^2
    double* partial;   // [grid][5][nbins]
};

enum { FIN_REDUCE = 1, FIN_SKEW = 2, FIN_MAXVEL = 4, FIN_DT = 8, FIN_ADVANCE = 16, FIN_EFOUR = 32, FIN_EPHYS = 64 };

struct FinArgs {
    DevScalars* scal;
    int flags;
    int nbins, nblk_epi, nblk_z;
    const double* hist_partial;
    const double* om_partial;
    const double* ed_partial;    // FIN_EFOUR: [nblk_epi] direct E partials
    const double* skew_partial;  // [nblk_z][2]; after a velocity pass [b][0] holds sum u.u (FIN_EPHYS)
    const unsigned long long* maxvel_bits;
    const unsigned long long* maxvel_sa_bits;
    double* ek;  // [nbins]
    int kfmin, kfmax;
    double force_eps;
    double Ax, Ay, Az, C_At, fixed_dt;
    int sa;
    double Ah, t1, phi[3], c[14], kk[2];
};

struct ShellArgs {
    cx<double>* uk[3];
    int nk1, nk2, nk3, nk3p;
    int64_t nmodes;  // padded
    int notcubic, Lx, Ly, Lz;
    const double* ft;
    int last_eff, nullify;
};

struct NormArgs {
    void* data;
    int Mx, My, Mz, Nx2, Ny2, Nz2;
    int notcubic, Lx, Ly, Lz, all;
    double nf;
};

struct RandArgs {
    cx<double>* uk[3];
    int nk1, nk2, nk3, nk3p;
    int64_t nmodes;  // padded
    int notcubic, Lx, Ly, Lz;
    uint64_t seed;
    double k0;
};

struct ClearArgs {
    unsigned long long* a;
    unsigned long long* b;
};

// A batched launch: one kernel over the slots of a shape class (runs with
// the same grid, precision and model).  Slot = blockIdx.y; its argument
// struct is args[slot] (args2[slot]: k_yepi's epilogue struct) in device
// memory; a slot whose steps-left counter left[slot] is 0 skips the launch,
// and the batched finalize decrements it.  The grid in x is the single run's,
// so a slot's arithmetic (and its reductions' order) is that of the run
// stepped alone.
struct BatchLaunch {
    const void* args;
    const void* args2;
    int* left;
    int nslots;
    const void* args3 = nullptr;  // k_yepi_inv: the slots' inverse y pass structs
};

constexpr int kMaxBatchSlots = 64;
struct SlotSteps {
    int32_t n[kMaxBatchSlots];  // steps each slot takes in the next batched chunk
};

// FFT lengths with batched kernel instances (the serving grids' 3/2-padded
// lengths)
constexpr bool batch_length(int L) { return L == 48 || L == 96 || L == 144 || L == 192 || L == 288; }

bool supported_length(int L);
hipError_t launch_init_random(const RandArgs& a, hipStream_t st);
hipError_t launch_scale_state(const ShellArgs& a, double sc, hipStream_t st);
int strided_tile(int L, int prec);
hipError_t launch_strided_fp64(int L, int dir, const StridedArgs& a, int A, hipStream_t st, const BatchLaunch* bl = nullptr);
hipError_t launch_strided_fp32(int L, int in_fp64, int dir, const StridedArgs& a, int A, hipStream_t st,
                               const BatchLaunch* bl = nullptr);
hipError_t launch_strided(int L, int prec, int in_fp64, int out_fp64, int dir, const StridedArgs& a, int A, hipStream_t st,
                          const BatchLaunch* bl = nullptr);
hipError_t launch_zpass(int L, int prec, int mode, const ZArgs& a, int grid, hipStream_t st, const BatchLaunch* bl = nullptr);
int zpass_grid(int L, int prec, const ZArgs& a, int64_t npen);  // < 0: no such length / HIP error
int zpass_kernel(int L, int prec, const ZArgs& a);              // kernel launch_zpass picks
hipError_t launch_zplain(int L, int prec, int dir, const ZPlainArgs& a, int grid, hipStream_t st);
// inv: the variant with the next step's inverse y pass (k_yepi_inv), which
// launch_yepi runs when ia (the inverse pass's arguments) is given; it sets
// *x1v
int yepi_grid(int L, int prec, int nbins, int nout, int64_t items, int max_grid, bool inv = false);
hipError_t launch_yepi(int L, int prec, const StridedArgs& a, int A, const EpiArgs& e, int grid, hipStream_t st,
                       const BatchLaunch* bl = nullptr, const StridedArgs* ia = nullptr, int* x1v = nullptr);
hipError_t launch_set_int(int* p, int v, hipStream_t st);
hipError_t launch_epilogue(int prec, const EpiArgs& a, int grid, hipStream_t st, const BatchLaunch* bl = nullptr);
hipError_t launch_finalize(const FinArgs& a, hipStream_t st, const BatchLaunch* bl = nullptr);
hipError_t launch_inv_spectra(int prec, const InvSpecArgs& a, int grid, double* out, hipStream_t st);
hipError_t launch_clear(unsigned long long* a, unsigned long long* b, hipStream_t st, const BatchLaunch* bl = nullptr);
hipError_t launch_set_left(int* left, const SlotSteps& s, int nslots, hipStream_t st);
hipError_t launch_stream_copy(const void* in, void* out, int64_t bytes, int blocks, hipStream_t st);
hipError_t launch_forcing(const double* ek, const double* forced, double* ft, int last_eff, const ShellArgs& a, hipStream_t st);
hipError_t launch_ccc(const ShellArgs& a, hipStream_t st);
hipError_t launch_normalize(int prec, const NormArgs& a, hipStream_t st);

}  // namespace spns
