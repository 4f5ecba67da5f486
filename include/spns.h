/*
 * spns.h -- C ABI of the MI355X pseudospectral HIT step.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * SPNS_E* code; spns_last_error() gives the message.  Calls on one context
 * are ordered on that context's HIP stream; use one host thread per context.
 *
 * The reference (adalbal/SpNS) has no FFI: its seam is the C++ class surface
 * `main` uses.  Each entry point below names the reference member it replaces.
 * The C++ facade spns_amd/csrc/hit.hpp re-exposes these under the
 * reference's own class and method names.
 *
 * Spectral state layout ("compact box", the reference's Fourier-dump order,
 * src/HIT_LES_output.cpp:46-53 with one rank): component-major
 * [c][i2][i1][i3] complex double (re, im), with
 *   k2 = i2 <= Ny/2 ? i2 : i2 - (2*(Ny/2)+1),  i2 in [0, 2*(Ny/2)+1)
 *   k1 = i1 <= Nx/2 ? i1 : i1 - (2*(Nx/2)+1),  i1 in [0, 2*(Nx/2)+1)
 *   k3 = i3,                                    i3 in [0, Nz/2]
 * i.e. the retained box |k1|<=Nx/2, |k2|<=Ny/2, 0<=k3<=Nz/2 of
 * src/HIT_LES.cpp:860 (modes masked by length factors are present and zero).
 */
#ifndef SPNS_H
#define SPNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SPNS_OK = 0,
    SPNS_EINVAL = 1,       /* bad argument */
    SPNS_EUNSUPPORTED = 2, /* grid length not compiled in, or feature off */
    SPNS_ENOMEM = 3,       /* device allocation failed */
    SPNS_EHIP = 4,         /* HIP runtime error */
    SPNS_ENONFINITE = 5,   /* the CFL reduction saw a NaN/Inf velocity */
    SPNS_ECHECK = 6        /* a self-check (spns_check_phys_four_energy) failed */
};

enum { SPNS_PREC_FP64 = 0, SPNS_PREC_FP32_SPLIT = 1 };

typedef struct spns_ctx spns_ctx;

/* Arguments of HIT::HIT (src/HIT_LES.h:23-29) after the L-factor expansion
 * of src/MAIN.cpp:216-253, plus the build's options. */
typedef struct spns_config {
    int32_t Nx, Ny, Nz;      /* retained modes (src/HIT_LES.h:23) */
    int32_t Mx, My, Mz;      /* FFT grid; 0 => (3N)/2 (src/MAIN.cpp:177) */
    double nu, C_Smag, C_At; /* C_Smag > 0 turns LES on (src/MAIN.cpp:272) */
    int32_t Lx, Ly, Lz;      /* length factors as HIT receives them */
    double omega_x, omega_y, omega_z;
    int32_t isNotCubic, isSelfAdaptive, isComplexConjugateCorrected;
    int32_t precision;       /* SPNS_PREC_FP64 | SPNS_PREC_FP32_SPLIT */
    double delta;            /* LES filter width; <= 0 => 2pi/cbrt(...) (src/HIT_LES.h:53) */
    double fixed_dt;         /* > 0 => fixed time step instead of the CFL step */
    int32_t force_on;        /* constant-power forcing (src/HIT_LES.cpp:243-278) */
    double force_eps;        /* injection rate */
    int32_t force_kmin, force_kmax;
    int32_t device;          /* HIP device ordinal */
    int32_t use_graph;       /* capture the step as a hipGraph */
    int32_t zpass_variant;   /* z-pass kernel: 0 default (staged inputs; DNS forward folded into the next
                                pencil's inverse), 1 block per pencil, 2 wave per pencil, 3 staged, not folded */
    int32_t gradient_invariants; /* keep the state at the start of each step (one more state copy per step)
                                    so spns_invariant_spectra can evaluate it */
    int32_t keep_step_start; /* keep it as well (implied by the two flags above) so that
                                spns_velocity_physical_step_start can give its physical field */
    int32_t reserved[3];
} spns_config;

typedef struct spns_scalars {
    double time;     /* HIT::gettime */
    double dt;       /* HIT::getAt */
    double nu;       /* HIT::getnu */
    double maxvel;   /* CFL velocity of the last antitransformed field */
    double ek;       /* HIT::getEk_Tot (E of the current state) */
    double om;       /* HIT::getOk_Tot (enstrophy; eps = nu*om) */
    double relambda; /* HIT::getReLambda */
    double skewness; /* <(du/dx)^3>/<(du/dx)^2>^1.5 at the start of the last step */
    double kappa;    /* self-adaptive scheme's kappa */
    int64_t iter;    /* steps taken */
} spns_scalars;

/* Replaces HIT::HIT + HIT_init allocation (src/HIT_LES.h:23-84, src/HIT_LES.cpp:770-1092). */
int spns_create(const spns_config *cfg, spns_ctx **out);
/* Reuses a context for another run of the same shape class (no reference
 * counterpart; the serving loop keeps finished runs' contexts instead of
 * freeing and allocating device memory, which synchronises the device): the
 * physics of cfg (nu, C_Smag, C_At, rotation, fixed_dt, forcing, delta) for
 * the same grid, precision, model and options; state, history, time and
 * iteration count are cleared.  Then set an initial field and call
 * spns_finish_init.  SPNS_EINVAL if anything but the physics differs. */
int spns_reconfigure(spns_ctx *ctx, const spns_config *cfg);
/* Replaces HIT::~HIT / HIT_destroy (src/HIT_LES.cpp:1094-1130). */
int spns_destroy(spns_ctx *ctx);
/* Device memory a context with this configuration allocates (bytes), for
 * admission by HBM footprint; -1 if the configuration is invalid. */
int64_t spns_device_bytes(const spns_config *cfg);
/* Number of compact modes per component, and of shells K=0..last_rad_max. */
int64_t spns_nmodes(const spns_ctx *ctx);
int32_t spns_nbins(const spns_ctx *ctx);
/* last_rad (src/HIT_LES.h:48), the K41 / forcing shell limit. */
int32_t spns_last_rad(const spns_ctx *ctx);
/* Human-readable message of the last failing call. */
const char *spns_last_error(const spns_ctx *ctx);

/* Initial fields.  Each leaves R = 0, time = 0, iter = 0; then call
 * spns_finish_init (src/HIT_LES.cpp:1037-1086). */
/* State in the compact layout above: 3*nmodes complex (6*nmodes doubles). */
int spns_set_state(spns_ctx *ctx, const double *state);
int spns_get_state(spns_ctx *ctx, double *state);
/* The state's i2 planes [i2_first, i2_first + n_i2) only: [c][n_i2][nk1][nk3]
 * complex double (the compact layout above restricted to those planes), for
 * checks on grids whose whole state does not fit the host comfortably. */
int spns_get_state_planes(spns_ctx *ctx, int32_t i2_first, int32_t n_i2, double *out);
/* R history (R of the last step), compact layout, as complex double.  The
 * FP32-split form holds it in FP32 on the device; it is widened here. */
int spns_get_rhist(spns_ctx *ctx, double *rhist);
/* HIT::Input_Real_Field (src/HIT_LES.cpp:542-599, 738-763): u,v,w are
 * [x][y][z] of nxf*nyf*nzf doubles; ek_file/ens_file receive
 * Ek_init_file / Enstrophy_init_file (src/HIT_LES.cpp:1268-1294). */
int spns_init_from_physical(spns_ctx *ctx, const double *u, const double *v, const double *w,
                            int32_t nxf, int32_t nyf, int32_t nzf, double *ek_file, double *ens_file);
/* HIT::Input_K41_Field (src/HIT_LES.cpp:502-540); nranks > 1 reproduces the
 * per-rank srand(0)/rand() draws of an nranks-slab MPI run. */
int spns_init_k41(spns_ctx *ctx, int32_t nranks);
/* Seeded random-phase solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2),
 * scaled to u_rms (no reference counterpart: synthetic workload input). */
int spns_init_random(spns_ctx *ctx, uint64_t seed, double k0, double urms);
/* Rest of HIT_init: R=0, antitransform, CFL step unless isReLambda, Ek_init. */
int spns_finish_init(spns_ctx *ctx, int32_t isReLambda);

/* HIT::New_Fractional_Step_Method (src/HIT_LES.cpp:9-25), nsteps times. */
int spns_step(spns_ctx *ctx, int32_t nsteps);
/* HIT::Recalculate_TimeStep (src/HIT_LES.cpp:365-371). */
int spns_recalculate_timestep(spns_ctx *ctx);
/* HIT::setnu (src/HIT_LES.h:282). */
int spns_set_nu(spns_ctx *ctx, double nu);
/* HIT::Forcing_Reynolds_Lambda (src/HIT_LES.cpp:485-489). */
int spns_forcing_reynolds_lambda(spns_ctx *ctx, double relambda);
/* HIT::Recalculate_Reynolds_Lambda (src/HIT_LES.cpp:478-483). */
int spns_recalculate_reynolds_lambda(spns_ctx *ctx, double *relambda);
/* HIT::Forcing_Energy_Cascade (src/HIT_LES.cpp:410-444); forced_ek has
 * last_input_rad+1 entries. */
int spns_forcing_energy_cascade(spns_ctx *ctx, const double *forced_ek, int64_t last_input_rad, int32_t nullify);
/* HIT::Complex_Conjugate_Correction (src/HIT_LES.cpp:446-476). */
int spns_complex_conjugate_correction(spns_ctx *ctx);
/* HIT::Recalculate_Energy_Cascade (src/HIT_LES.cpp:1261-1266): E(K) for
 * K = 0..nbins-1 into ek (may be NULL), total into ek_tot (may be NULL). */
int spns_energy_cascade(spns_ctx *ctx, double *ek, int32_t nbins, double *ek_tot);
/* The kernels this context's step runs (no reference counterpart;
 * diagnostics and tests): *zpass_kernel = 2 staged z pass, 1 block per
 * pencil, 0 wave per pencil; *fused_epilogue = 1 when the DNS step fuses its
 * forward y pass with the epilogue (one kernel), 2 when that kernel also
 * runs the next step's inverse y pass (the step's own inverse y pass then
 * does nothing), else 0.  Either may be NULL. */
int spns_pass_kernels(spns_ctx *ctx, int32_t *zpass_kernel, int32_t *fused_epilogue);
/* HIT getters (src/HIT_LES.h:187-201, 277-291). */
int spns_get_scalars(spns_ctx *ctx, spns_scalars *out);
/* HIT::Recalculate_Velocity_Antitransform + RealVelocity_to_BinaryFile's
 * data (src/HIT_LES.cpp:60-69): u,v,w as [3][Mx][My][Mz] doubles. */
int spns_velocity_physical(spns_ctx *ctx, double *uvw);
/* RealVelocity_to_BinaryFile's data between steps (src/HIT_LES_output.cpp:12-14):
 * the reference's u,v,w arrays then hold the physical field of the state at
 * the start of the last step (src/HIT_LES.cpp:9-25; the current state before
 * any step).  Same layout as spns_velocity_physical; the CFL maxima are left
 * alone.  Needs a kept step-start state (keep_step_start, gradient_invariants
 * or isSelfAdaptive) once a step has run; SPNS_EUNSUPPORTED otherwise. */
int spns_velocity_physical_step_start(spns_ctx *ctx, double *uvw);
/* HIT::Check_PhysFour_Energy (src/HIT_LES.cpp:1216-1252), the QA build's
 * check at the start of every step: the energy of the current state summed
 * directly over the retained modes (ek_four), by shells (ek_spns,
 * HIT::Recalculate_Energy) and in physical space (ek_real = 1/2 sum u.u /
 * (Mx My Mz) over the FFT grid, FP64 sums).  Returns SPNS_ECHECK, with the
 * reference's message, if |ek_real - ek_four| or |ek_spns - ek_four| exceeds
 * atol (the reference crashes there).  Leaves the step's scalars, spectrum
 * and CFL maxima unchanged.  Outputs may be NULL. */
int spns_check_phys_four_energy(spns_ctx *ctx, double atol, double *ek_real, double *ek_four, double *ek_spns);
/* HIT::Recalculate_Velocity_Antitransform (src/HIT_LES.cpp:60-69) without
 * copying the field out: refreshes the CFL maxima from the current state. */
int spns_recalculate_velocity_antitransform(spns_ctx *ctx);
/* HIT::Recalculate_Invariants_Distribution (src/HIT_LES.cpp:1297-1319) without
 * the files: shell spectra, K = 0..spns_nbins-1, of the velocity-gradient
 * invariants QG, QS, RG, RS, V2 (src/tensor_algebra.cpp:250-303) into
 * spectra[f * nbins + K], f in that order; nbins >= spns_nbins(ctx).  Like
 * the reference, of the gradients of the field at the start of the last step
 * (the current field before any step).  Needs config.gradient_invariants;
 * SPNS_EUNSUPPORTED otherwise. */
int spns_invariant_spectra(spns_ctx *ctx, int32_t nbins, double *spectra);
/* spns_invariant_spectra without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops with a fixed footprint, harnesses that
 * forbid hipMalloc).  The shell sums' block partials and results live in a
 * fixed 4 MiB device array each GPU holds for the life of the process: the
 * call needs (min(1024, nk1*nk2) + 1) * 5 * nbins doubles of it, at most
 * 2^19 (up to 102 shells on larger cross-sections, cubic grids up to
 * N = 116); larger grids return SPNS_EINVAL.  Same results as
 * spns_invariant_spectra, bit for bit.  All calls on one GPU share that
 * array: calls that overlap on one GPU (several host threads) race. */
int spns_invariant_spectra_noalloc(spns_ctx *ctx, int32_t nbins, double *spectra);
/* Wait for the context's stream. */
int spns_synchronize(spns_ctx *ctx);
/* Non-blocking: *done = 1 when every step enqueued by the last spns_step call
 * has finished on the device, else 0.  Once done, returns SPNS_ENONFINITE
 * (with *done = 1) if the run diverged: its CFL velocity, energy, enstrophy
 * or time step is NaN/Inf.  No reference counterpart (the serving loop polls
 * its runs with it and retires a diverged run alone, where the reference's
 * crash() would abort every rank, src/util.cpp:93-116). */
int spns_query(spns_ctx *ctx, int32_t *done);
/* Measured practical HBM ceiling: a streaming copy kernel (16-byte
 * non-temporal loads and stores) between two device buffers of `bytes`
 * each, reps times; *gb_per_s = bytes read + written per second (1e9 B).
 * No reference counterpart (measurement). */
int spns_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double *gb_per_s);
/* Free and total device memory of a HIP device (hipMemGetInfo), for
 * admission by HBM footprint.  No reference counterpart. */
int spns_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes);
/* Device time (ms) of the last spns_step call's kernels, from HIP events. */
int spns_last_step_ms(spns_ctx *ctx, double *ms);
/* Runs nsteps steps launched pass by pass with HIP events between the passes
 * on the context's stream and returns the mean ms of each pass in ms[0..7]:
 * clear, inv_y, inv_x, z pass, fwd_x, fwd_y, epilogue, finalize.  The steps
 * advance the solution like spns_step (no reference counterpart). */
int spns_profile_passes(spns_ctx *ctx, int32_t nsteps, double *ms);

/* ---- Batched stepping of several runs (no reference counterpart) ----
 * The serving loop advances independent runs of one shape class (same N, M,
 * precision and model: DNS or LES) together: each pass of the step is one
 * kernel launch over every slot of the batch, and a chunk of steps is one
 * captured hipGraph.  Slots are ordinary contexts (spns_create + an initial
 * field + spns_finish_init), attached while the batch is idle; their
 * physics parameters (nu, C_s, rotation, forcing, time step) may differ.
 * Each slot computes bit for bit what spns_step would on its own.  A
 * context attached to a batch must not be stepped with spns_step, and its
 * other calls must wait until the batch is idle (spns_batch_query reports
 * done, or spns_batch_synchronize).  Grids whose 3/2-padded FFT lengths are
 * 48, 96, 144, 192 or 288 have batched kernels; others return
 * SPNS_EUNSUPPORTED at attach. */
typedef struct spns_batch spns_batch;
/* nslots in 1..64; chunk = the most steps one spns_batch_step takes. */
int spns_batch_create(int32_t device, int32_t nslots, int32_t chunk, spns_batch **out);
int spns_batch_destroy(spns_batch *b);
int spns_batch_attach(spns_batch *b, int32_t slot, spns_ctx *ctx);
int spns_batch_detach(spns_batch *b, int32_t slot);
/* Advances slot s by nsteps[s] steps (0..chunk; 0 or an empty slot idles),
 * asynchronously, on the batch's stream. */
int spns_batch_step(spns_batch *b, const int32_t *nsteps);
/* Non-blocking: *done = 1 when the last spns_batch_step has finished.  Then
 * spns_query on a slot's context reports SPNS_ENONFINITE if it diverged. */
int spns_batch_query(spns_batch *b, int32_t *done);
int spns_batch_synchronize(spns_batch *b);
/* A slot's scalars after the last spns_batch_step, from a copy the batch
 * makes on its own stream after every chunk (no device read here); once
 * spns_batch_query has reported that chunk done (fetches queued after it do
 * not invalidate them).  SPNS_ENONFINITE if the run diverged (the scalars
 * are still written). */
int spns_batch_slot_scalars(spns_batch *b, int32_t slot, spns_scalars *out);
/* Queues copies of a slot's state (compact layout, 6*nmodes doubles) and
 * spectrum (spns_nbins doubles) into host buffers on the batch's stream
 * (either may be NULL); they are complete when spns_batch_query reports
 * done.  With pinned buffers (spns_host_alloc) the copies do not block the
 * caller; the slot must not be stepped until then. */
int spns_batch_fetch(spns_batch *b, int32_t slot, double *state, double *ek);
const char *spns_batch_last_error(const spns_batch *b);
/* Pinned host memory (hipHostMalloc) for asynchronous copies. */
int spns_host_alloc(int64_t bytes, void **out);
int spns_host_free(void *p);

/* ---- FFT_package drop-in (src/FFTW_package.h:21-71) on device buffers ----
 * complex layout [y][x][Mz/2+1] (FFTW_MPI_TRANSPOSED, one rank),
 * real layout [x][y][2(Mz/2+1)] (src/util.h:75-80).  Both transforms are
 * unnormalised and destroy their input, like the FFTW plans they replace.
 * Ordering: like fftw_execute, spns_fft_r2c, spns_fft_c2r and
 * spns_fft_normalize return with their output complete.  They run on a
 * blocking HIP stream of their own, so work the caller queued on the legacy
 * default stream before the call (filling the input, zeroing the output) is
 * finished before they start; work on the caller's non-blocking streams must
 * be synchronised by the caller first. */
typedef struct spns_fft spns_fft;
/* FFT_package::FFT_package / FFT_init (src/FFTW_package.cpp:47-56). */
int spns_fft_create(int32_t Mx, int32_t My, int32_t Mz, int32_t precision, int32_t device, spns_fft **out);
int spns_fft_destroy(spns_fft *fft);
/* FFT_package::FFTWr2c (src/FFTW_package.cpp:7-9). */
int spns_fft_r2c(spns_fft *fft, void *rin, void *cout);
/* FFT_package::IFFTWc2r (src/FFTW_package.cpp:12-14). */
int spns_fft_c2r(spns_fft *fft, void *cin, void *rout);
/* FFT_package::Fourier_Normalization(COMPLEX*, relevant) (src/FFTW_package.cpp:33-42)
 * with relevant = the retained box of (Nx,Ny,Nz); Nx<=0 normalises all. */
int spns_fft_normalize(spns_fft *fft, void *c, int32_t Nx, int32_t Ny, int32_t Nz);
int spns_fft_synchronize(spns_fft *fft);
const char *spns_fft_last_error(const spns_fft *fft);

#ifdef __cplusplus
}
#endif

#endif /* SPNS_H */
