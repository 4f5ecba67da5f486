/*
 * spns_oracle.c -- FP64 CPU restatement of SpNS's pseudospectral HIT time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path (spns_amd/).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it; the product never links or calls it.
 *
 * It restates, one MPI rank, the reference algorithm:
 *   - data layouts of src/util.h:66-80 (LOOP_FOURIER [b=y][a=x][k3],
 *     LOOP_REAL [x][y][z] padded to 2(Mz/2+1)), FOURIER_FREQ src/util.h:67
 *   - masks rad2/dealiased/conjugate  src/HIT_LES.cpp:833-881
 *   - c2r/r2c + normalisation         src/FFTW_package.cpp:7-42 (FFTW itself
 *     is replaced by the Stockham FFT below; FFTW is not in the image)
 *   - the step                         src/HIT_LES.cpp:9-403
 *   - constant-power forcing (#if 0)   src/HIT_LES.cpp:243-278 (opt-in)
 *   - forcings                         src/HIT_LES.cpp:410-495
 *   - K41 / file initial fields        src/HIT_LES.cpp:502-763
 *   - shell reductions                 src/HIT_LES.cpp:1137-1266
 *   - file spectrum energy             src/HIT_LES.cpp:1268-1294
 * Each function below names the lines it follows.
 *
 * Pinned by tests/test_oracle_*.py: FFTs against numpy.fft and brute-force
 * DFT; Parseval (the reference's own QA invariant, src/HIT_LES.cpp:1216);
 * closed forms (single modes, viscous AB2 recurrence, projection); and the
 * reference's QA REF logs (QA/TEST_1, QA/TEST_2 per-step Time/Dt/Ek/Om/ReLambda,
 * and QA/TEST_3..9 with 4-rank K41 emulation).
 *
 * Extensions beyond the reference (documented; the reference's own tests do
 * not cover them): fixed time step, explicit LES Delta, constant-power
 * forcing switch, velocity-gradient skewness, the seeded synthetic initial
 * field.  Each is pinned by tests/test_oracle_*.py instead: closed forms for
 * the fixed step and the forcing law, a numpy restatement and closed forms
 * for the skewness sums, SplitMix64's published outputs and the field's
 * exact per-mode amplitude, solenoidality and Hermitian plane for the
 * initial field.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643 /* src/util.h:57 */
#define PB 16 /* pencils gathered together in the strided passes */

/* ------------------------------------------------------------------------ */
/* Twiddles: exp(-2 pi i m/n) with octant reduction in long double so that   */
/* quarter/eighth points are exact by symmetry.                              */
/* ------------------------------------------------------------------------ */
static void twiddle_fwd(long m, long n, double *re, double *im)
{
    m %= n;
    if (m < 0) m += n;
    long a = 8 * m; /* t = a/(8n) in [0,1) */
    int sflip = 0, cflip = 0, swap = 0;
    if (a > 4 * n) { a = 8 * n - a; sflip = 1; }
    if (a > 2 * n) { a = 4 * n - a; cflip = 1; }
    if (a > n)     { a = 2 * n - a; swap = 1; }
    long double th = (long double)ORC_PI * (long double)a / (4.0L * (long double)n);
    long double c = cosl(th), s = sinl(th);
    if (a == 0) { c = 1.0L; s = 0.0L; }
    if (swap) { long double t = c; c = s; s = t; }
    if (cflip) c = -c;
    if (sflip) s = -s;
    *re = (double)c;
    *im = (double)(-s);
}

/* ------------------------------------------------------------------------ */
/* 1-D complex FFT: Stockham autosort, radices 4,2,3,5 (textbook form, e.g.  */
/* Govindaraju et al., SC08).  Unnormalised.  sign=-1 forward, +1 backward.  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, nfac;
    int fac[64];
    double *tw; /* 2n: forward twiddles exp(-2 pi i m/n) */
} plan1d;

static int plan1d_init(plan1d *p, int n)
{
    p->n = n;
    p->nfac = 0;
    int r = n;
    while (r % 4 == 0) { p->fac[p->nfac++] = 4; r /= 4; }
    while (r % 2 == 0) { p->fac[p->nfac++] = 2; r /= 2; }
    while (r % 3 == 0) { p->fac[p->nfac++] = 3; r /= 3; }
    while (r % 5 == 0) { p->fac[p->nfac++] = 5; r /= 5; }
    if (r != 1) return -1;
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int m = 0; m < n; m++) twiddle_fwd(m, n, &p->tw[2 * m], &p->tw[2 * m + 1]);
    return 0;
}

static void plan1d_free(plan1d *p) { free(p->tw); p->tw = NULL; }

/* x, w: 2n doubles each; result lands in x. */
static void fft1d(const plan1d *p, double *x, double *w, int sign)
{
    const int n = p->n;
    double *src = x, *dst = w;
    int Ns = 1;
    for (int f = 0; f < p->nfac; f++) {
        const int R = p->fac[f];
        const int nR = n / R;
        const int tstep = n / (Ns * R);
        for (int j = 0; j < nR; j++) {
            const int k = j % Ns;
            double ar[5], ai[5];
            for (int r = 0; r < R; r++) {
                double xr = src[2 * (j + r * nR)], xi = src[2 * (j + r * nR) + 1];
                const int t = r * k * tstep;
                double c = p->tw[2 * t], s = p->tw[2 * t + 1];
                if (sign > 0) s = -s;
                ar[r] = xr * c - xi * s;
                ai[r] = xr * s + xi * c;
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int q = 0; q < R; q++) {
                double br = 0.0, bi = 0.0;
                for (int r = 0; r < R; r++) {
                    const int t = ((r * q) % R) * nR;
                    double c = p->tw[2 * t], s = p->tw[2 * t + 1];
                    if (sign > 0) s = -s;
                    br += ar[r] * c - ai[r] * s;
                    bi += ar[r] * s + ai[r] * c;
                }
                dst[2 * (base + q * Ns)] = br;
                dst[2 * (base + q * Ns) + 1] = bi;
            }
        }
        double *t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* 3-D transforms in the reference's single-rank layouts                     */
/* (src/FFTW_package.cpp:47-56 with FFTW_MPI_TRANSPOSED_OUT/IN):             */
/*   complex  [b: y][a: x][k3 <= Mz/2]                                       */
/*   real     [i: x][j: y][k: z] padded to 2(Mz/2+1)                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Mx, My, Mz, Mz2;
    plan1d px, py, pz;
    int nthreads;
    double *scratch; /* per-thread: 2 (PB+1) max(M) doubles */
    int maxM;
} fft3d;

static int fft3d_init(fft3d *f, int Mx, int My, int Mz)
{
    f->Mx = Mx; f->My = My; f->Mz = Mz; f->Mz2 = Mz / 2;
    if (plan1d_init(&f->px, Mx) || plan1d_init(&f->py, My) || plan1d_init(&f->pz, Mz)) return -1;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    f->nthreads = nt;
    f->maxM = Mx > My ? Mx : My;
    if (Mz > f->maxM) f->maxM = Mz;
    f->scratch = (double *)malloc(sizeof(double) * 2 * (PB + 1) * (size_t)f->maxM * nt);
    return 0;
}

static void fft3d_free(fft3d *f)
{
    plan1d_free(&f->px); plan1d_free(&f->py); plan1d_free(&f->pz);
    free(f->scratch);
}

static double *thread_scratch(const fft3d *f)
{
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    return f->scratch + (size_t)2 * (PB + 1) * f->maxM * t;
}

/* in-place complex transform along x and y of a [My][Mx][Mz2+1] array.
 * Pencils are gathered in groups of up to PB consecutive k3 (or x) so each
 * read touches whole cache lines. */
static void fft_xy(const fft3d *f, double *c, int sign)
{
    const int Mx = f->Mx, My = f->My, nz = f->Mz2 + 1;
    const int ngz = (nz + PB - 1) / PB;
    /* x: for each (b, k3 group), pencils over a, stride nz */
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < My; b++)
        for (int g = 0; g < ngz; g++) {
            double *buf = thread_scratch(f), *w = buf + 2 * (size_t)f->maxM * PB;
            const int k0 = g * PB, nb = (nz - k0 < PB) ? nz - k0 : PB;
            for (int a = 0; a < Mx; a++) {
                const double *src = c + 2 * (((size_t)b * Mx + a) * nz + k0);
                for (int q = 0; q < nb; q++) {
                    buf[2 * ((size_t)q * Mx + a)] = src[2 * q];
                    buf[2 * ((size_t)q * Mx + a) + 1] = src[2 * q + 1];
                }
            }
            for (int q = 0; q < nb; q++) fft1d(&f->px, buf + 2 * (size_t)q * Mx, w, sign);
            for (int a = 0; a < Mx; a++) {
                double *dst = c + 2 * (((size_t)b * Mx + a) * nz + k0);
                for (int q = 0; q < nb; q++) {
                    dst[2 * q] = buf[2 * ((size_t)q * Mx + a)];
                    dst[2 * q + 1] = buf[2 * ((size_t)q * Mx + a) + 1];
                }
            }
        }
    /* y: for each (a,k3) column group, pencils over b, stride Mx*nz */
    const size_t row = (size_t)Mx * nz;
    const size_t ncol = row, ngc = (ncol + PB - 1) / PB;
#pragma omp parallel for schedule(static)
    for (size_t g = 0; g < ngc; g++) {
        double *buf = thread_scratch(f), *w = buf + 2 * (size_t)f->maxM * PB;
        const size_t c0 = g * PB;
        const int nb = (int)((ncol - c0 < PB) ? ncol - c0 : PB);
        for (int b = 0; b < My; b++) {
            const double *src = c + 2 * ((size_t)b * row + c0);
            for (int q = 0; q < nb; q++) {
                buf[2 * ((size_t)q * My + b)] = src[2 * q];
                buf[2 * ((size_t)q * My + b) + 1] = src[2 * q + 1];
            }
        }
        for (int q = 0; q < nb; q++) fft1d(&f->py, buf + 2 * (size_t)q * My, w, sign);
        for (int b = 0; b < My; b++) {
            double *dst = c + 2 * ((size_t)b * row + c0);
            for (int q = 0; q < nb; q++) {
                dst[2 * q] = buf[2 * ((size_t)q * My + b)];
                dst[2 * q + 1] = buf[2 * ((size_t)q * My + b) + 1];
            }
        }
    }
}

/* c2r: complex cin [My][Mx][nz] -> real rout [Mx][My][2nz]; cin preserved.
 * work: complex array of the same size as cin.  The real (z) dimension is
 * done last and ignores Im of the k3=0 and k3=Mz/2 entries, as a c2r does. */
static void c2r3d(const fft3d *f, const double *cin, double *rout, double *work)
{
    const int Mx = f->Mx, My = f->My, Mz = f->Mz, nz = f->Mz2 + 1;
    memcpy(work, cin, sizeof(double) * 2 * (size_t)My * Mx * nz);
    fft_xy(f, work, +1);
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++) {
            double *buf = thread_scratch(f), *w = buf + 2 * (size_t)f->maxM * PB;
            const double *h = work + 2 * (((size_t)b * Mx + a) * nz);
            for (int k = 0; k < Mz; k++) { buf[2 * k] = 0.0; buf[2 * k + 1] = 0.0; }
            for (int k = 0; k < nz; k++) {
                buf[2 * k] = h[2 * k];
                buf[2 * k + 1] = h[2 * k + 1];
                if (k > 0 && Mz - k != k) {
                    buf[2 * (Mz - k)] = h[2 * k];
                    buf[2 * (Mz - k) + 1] = -h[2 * k + 1];
                }
            }
            buf[1] = 0.0;
            if (Mz % 2 == 0) buf[2 * (Mz / 2) + 1] = 0.0;
            fft1d(&f->pz, buf, w, +1);
            double *o = rout + ((size_t)a * My + b) * (2 * (size_t)nz);
            for (int k = 0; k < Mz; k++) o[k] = buf[2 * k];
        }
}

/* r2c: real rin [Mx][My][2nz] -> complex cout [My][Mx][nz] (unnormalised) */
static void r2c3d(const fft3d *f, const double *rin, double *cout)
{
    const int Mx = f->Mx, My = f->My, Mz = f->Mz, nz = f->Mz2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++) {
            double *buf = thread_scratch(f), *w = buf + 2 * (size_t)f->maxM * PB;
            const double *x = rin + ((size_t)i * My + j) * (2 * (size_t)nz);
            for (int k = 0; k < Mz; k++) { buf[2 * k] = x[k]; buf[2 * k + 1] = 0.0; }
            fft1d(&f->pz, buf, w, -1);
            double *o = cout + 2 * (((size_t)j * Mx + i) * nz);
            for (int k = 0; k < nz; k++) { o[2 * k] = buf[2 * k]; o[2 * k + 1] = buf[2 * k + 1]; }
        }
    fft_xy(f, cout, -1);
}

/* ------------------------------------------------------------------------ */
/* Solver state                                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    /* HIT constructor arguments (src/HIT_LES.h:23-29), after MAIN's L-factor
     * expansion (src/MAIN.cpp:216-253). */
    int Nx, Ny, Nz, Mx, My, Mz;
    double nu, C_Smag, C_At;
    int Lx, Ly, Lz;
    double omega_x, omega_y, omega_z;
    int isLES, isRotating, isNotCubic, isSelfAdaptive, isCCC;
    /* extensions */
    double delta;    /* <= 0: reference formula src/HIT_LES.h:53 */
    double fixed_dt; /* > 0: fixed time step instead of CFL */
    int force_on;    /* constant-power forcing src/HIT_LES.cpp:243-278 */
    double force_eps;
    int force_kmin, force_kmax;
} orc_params;

typedef struct {
    orc_params p;
    int ActualMx, ActualMy, ActualMz;
    int Nx_2, Ny_2, Nz_2, Mz2, last_rad, last_rad_max;
    double delta, Ax, Ay, Az;
    size_t nc, nr; /* complex / real element counts */
    fft3d fft;
    /* state */
    double *uk[3], *uk0[3], *ukaux[3], *R0[3], *R1[3], *conv_k[3];
    double *uk_start[3]; /* state at the start of the last step (orc_keep_start) */
    double *u[3], *grad[9], *conv, *ctmp;
    double *nu_eddy, *sgs[6], *div_k[3];
    unsigned char *dealiased, *conjugate;
    double *rad2;
    double *Ek;
    /* scalars */
    double time, At, MaxVel, C_visc, C_conv;
    double Ah, t1, kappa, kk[2], phi[3], cc[14], phi0, kappa05;
    double Ek_Tot, Ok_Tot, ReLambda, Ek_init;
    double skew_sum2, skew_sum3; /* of du/dx over the M grid, last step */
    long iter;                   /* src/HIT_LES.cpp:3 */
} orc_t;

static inline int ffreq(int a, int M) { return (a <= M / 2) ? a : a - M; } /* src/util.h:67 */

#define CIX(o, b, a, k3) ((((size_t)(b) * (o)->p.Mx + (a)) * ((o)->Mz2 + 1)) + (k3))

static double *calloc_d(size_t n) { return (double *)calloc(n, sizeof(double)); }

static inline int shell_of(int k2) /* K = round(sqrt(k2)), src/HIT_LES.cpp:1166 */
{
    return (int)lround(sqrt((double)k2));
}

/* src/HIT_LES.cpp:833-881 */
static void init_masks(orc_t *o)
{
    const int Mx = o->p.Mx, My = o->p.My;
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 <= o->Mz2; k3++) {
                size_t ind = CIX(o, b, a, k3);
                int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
                o->rad2[ind] = (double)(k1 * k1 + k2 * k2 + k3 * k3);
                int d = (abs(k1) <= o->Nx_2 && abs(k2) <= o->Ny_2 && k3 <= o->Nz_2);
                if (o->p.isNotCubic) {
                    if ((k1 % o->p.Lx != 0) || (k2 % o->p.Ly != 0) || (k3 % o->p.Lz != 0)) d = 0;
                }
                o->dealiased[ind] = (unsigned char)d;
                o->conjugate[ind] = (unsigned char)(d && (k3 != 0) && !((o->p.Mz % 2 == 0) && (k3 == o->Mz2)));
            }
}

orc_t *orc_create(const orc_params *p)
{
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    o->p = *p;
    /* src/HIT_LES.h:31-58 */
    o->ActualMx = (p->Mx % 2 == 0) ? p->Mx / p->Lx : (p->Mx - 1) / p->Lx + 1;
    o->ActualMy = (p->My % 2 == 0) ? p->My / p->Ly : (p->My - 1) / p->Ly + 1;
    o->ActualMz = (p->Mz % 2 == 0) ? p->Mz / p->Lz : (p->Mz - 1) / p->Lz + 1;
    o->Nx_2 = p->Nx / 2; o->Ny_2 = p->Ny / 2; o->Nz_2 = p->Nz / 2;
    o->Mz2 = p->Mz / 2;
    int lr = (p->Nx - 1) / 2;
    if ((p->Ny - 1) / 2 < lr) lr = (p->Ny - 1) / 2;
    if ((p->Nz - 1) / 2 < lr) lr = (p->Nz - 1) / 2;
    o->last_rad = lr;
    o->last_rad_max = (int)ceil(sqrt(1.0 * o->Nx_2 * o->Nx_2 + 1.0 * o->Ny_2 * o->Ny_2 + 1.0 * o->Nz_2 * o->Nz_2));
    o->delta = (p->delta > 0.0) ? p->delta
        : 2.0 * ORC_PI / cbrt((double)(p->Lx * p->Ly * p->Lz) * o->ActualMx * o->ActualMy * o->ActualMz);
    o->Ax = 2.0 * ORC_PI / o->ActualMx / p->Lx;
    o->Ay = 2.0 * ORC_PI / o->ActualMy / p->Ly;
    o->Az = 2.0 * ORC_PI / o->ActualMz / p->Lz;
    if (fft3d_init(&o->fft, p->Mx, p->My, p->Mz)) { free(o); return NULL; }
    o->nc = (size_t)p->My * p->Mx * (o->Mz2 + 1);
    o->nr = (size_t)p->Mx * p->My * 2 * (o->Mz2 + 1);
    for (int c = 0; c < 3; c++) {
        o->uk[c] = calloc_d(2 * o->nc);
        o->R0[c] = calloc_d(2 * o->nc);
        o->R1[c] = calloc_d(2 * o->nc);
        o->conv_k[c] = calloc_d(2 * o->nc);
        o->u[c] = calloc_d(o->nr);
        if (p->isSelfAdaptive) { o->uk0[c] = calloc_d(2 * o->nc); o->ukaux[c] = calloc_d(2 * o->nc); }
        if (p->isLES) o->div_k[c] = calloc_d(2 * o->nc);
    }
    int ngrad = p->isLES ? 9 : 3;
    for (int g = 0; g < ngrad; g++) o->grad[g] = calloc_d(o->nr);
    o->conv = calloc_d(o->nr);
    o->ctmp = calloc_d(2 * o->nc);
    if (p->isLES) {
        o->nu_eddy = calloc_d(o->nr);
        for (int s = 0; s < 6; s++) o->sgs[s] = calloc_d(o->nr);
    }
    o->dealiased = (unsigned char *)calloc(o->nc, 1);
    o->conjugate = (unsigned char *)calloc(o->nc, 1);
    o->rad2 = calloc_d(o->nc);
    o->Ek = calloc_d((size_t)o->last_rad_max + 1);
    init_masks(o);
    /* self-adaptive constants src/HIT_LES.cpp:1049-1078 */
    if (p->isSelfAdaptive) {
        o->Ah = (o->Ax + o->Ay + o->Az) / 3.0;
        o->t1 = 0.9302468;
        o->kappa = 0.5;
        o->kk[0] = 0.73782212; o->kk[1] = 0.44660387;
        o->phi[0] = atan(164.0 / 99.0); o->phi[1] = ORC_PI / 3.0; o->phi[2] = ORC_PI * 3.0 * 3.0 / (5.0 * 5.0);
        const double cc[14] = {0.0647998, -0.386022, 3.72945, -9.38143, 7.06574, 2403400.0, -5018490.0,
                               2620140.0, 2945.0, -6665.76, 3790.54, 4.80513, -16.9473, 15.0155};
        memcpy(o->cc, cc, sizeof(cc));
    }
    return o;
}

void orc_destroy(orc_t *o)
{
    if (!o) return;
    for (int c = 0; c < 3; c++) {
        free(o->uk[c]); free(o->R0[c]); free(o->R1[c]); free(o->conv_k[c]); free(o->u[c]);
        free(o->uk0[c]); free(o->ukaux[c]); free(o->div_k[c]); free(o->uk_start[c]);
    }
    for (int g = 0; g < 9; g++) free(o->grad[g]);
    free(o->conv); free(o->ctmp); free(o->nu_eddy);
    for (int s = 0; s < 6; s++) free(o->sgs[s]);
    free(o->dealiased); free(o->conjugate); free(o->rad2); free(o->Ek);
    fft3d_free(&o->fft);
    free(o);
}

/* Touches every page of the field arrays (writing the zeros calloc gave
   them), in parallel, so that a timed step does not pay the first-touch
   page faults of the large grids.  Values are unchanged. */
static void prefault_d(double *a, size_t n)
{
    if (!a) return;
    const size_t stride = 4096 / sizeof(double);
    const long npages = (long)((n + stride - 1) / stride);
#pragma omp parallel for schedule(static)
    for (long pg = 0; pg < npages; pg++) {
        volatile double *q = a + (size_t)pg * stride;
        *q = *q;
    }
}

void orc_prefault(orc_t *o)
{
    for (int c = 0; c < 3; c++) {
        prefault_d(o->uk[c], 2 * o->nc); prefault_d(o->R0[c], 2 * o->nc); prefault_d(o->R1[c], 2 * o->nc);
        prefault_d(o->conv_k[c], 2 * o->nc); prefault_d(o->u[c], o->nr);
        prefault_d(o->uk0[c], 2 * o->nc); prefault_d(o->ukaux[c], 2 * o->nc); prefault_d(o->div_k[c], 2 * o->nc);
    }
    for (int g = 0; g < 9; g++) prefault_d(o->grad[g], o->nr);
    prefault_d(o->conv, o->nr);
    prefault_d(o->ctmp, 2 * o->nc);
    prefault_d(o->nu_eddy, o->nr);
    for (int s = 0; s < 6; s++) prefault_d(o->sgs[s], o->nr);
}

/* ------------------------------------------------------------------------ */
/* Normalisation on D (src/FFTW_package.cpp:33-42)                           */
static void normalize_D(orc_t *o, double *ck)
{
    const double nf = (double)((long)o->p.Mx * o->p.My * o->p.Mz);
#pragma omp parallel for schedule(static)
    for (size_t ind = 0; ind < o->nc; ind++)
        if (o->dealiased[ind]) { ck[2 * ind] /= nf; ck[2 * ind + 1] /= nf; }
}

/* i k_dir * uk on D, 0 off D (src/HIT_LES.cpp:34-58) */
static void gradient_spectrum(orc_t *o, const double *uk, int dir, double *out)
{
    const int Mx = o->p.Mx, My = o->p.My;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 <= o->Mz2; k3++) {
                size_t ind = CIX(o, b, a, k3);
                int kd = (dir == 0) ? ffreq(a, Mx) : (dir == 1) ? ffreq(b, My) : k3;
                if (o->dealiased[ind]) {
                    out[2 * ind] = -kd * uk[2 * ind + 1];
                    out[2 * ind + 1] = kd * uk[2 * ind];
                } else {
                    out[2 * ind] = 0.0; out[2 * ind + 1] = 0.0;
                }
            }
}

/* src/HIT_LES.cpp:60-69 (non-destructive c2r) */
void orc_antitransform(orc_t *o)
{
    for (int c = 0; c < 3; c++) c2r3d(&o->fft, o->uk[c], o->u[c], o->ctmp);
}

static void real_loop_product(orc_t *o, double *out, const double *g0, const double *g1, const double *g2)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    const double *u = o->u[0], *v = o->u[1], *w = o->u[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++)
            for (int k = 0; k < Mz; k++) {
                size_t ind = ((size_t)i * My + j) * 2 * nz + k;
                out[ind] = u[ind] * g0[ind] + v[ind] * g1[ind] + w[ind] * g2[ind]; /* :119 */
            }
}

static void skew_accumulate(orc_t *o, const double *g)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    long double s2 = 0.0L, s3 = 0.0L; /* long double: see integrate_field */
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : s2, s3)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++)
            for (int k = 0; k < Mz; k++) {
                double x = g[((size_t)i * My + j) * 2 * nz + k];
                s2 += x * x;
                s3 += x * x * x;
            }
    o->skew_sum2 = (double)s2;
    o->skew_sum3 = (double)s3;
}

/* Convective term, advective form (src/HIT_LES.cpp:74-125).  For DNS the
 * three gradient fields of one velocity component are formed at a time; for
 * LES all nine are kept for the eddy viscosity. */
static void convective(orc_t *o)
{
    orc_antitransform(o);
    for (int c = 0; c < 3; c++) {
        double *g[3];
        for (int d = 0; d < 3; d++) {
            g[d] = o->p.isLES ? o->grad[3 * c + d] : o->grad[d];
            gradient_spectrum(o, o->uk[c], d, o->ctmp);
            c2r3d(&o->fft, o->ctmp, g[d], o->conv_k[c]); /* conv_k[c] as scratch */
        }
        if (c == 0) skew_accumulate(o, g[0]);
        real_loop_product(o, o->conv, g[0], g[1], g[2]);
        r2c3d(&o->fft, o->conv, o->conv_k[c]);
        normalize_D(o, o->conv_k[c]);
    }
}

/* LES: src/HIT_LES.cpp:130-192 */
static void sgs_divergence(orc_t *o)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    double *const *gu = &o->grad[0], *const *gv = &o->grad[3], *const *gw = &o->grad[6];
    const double ct = o->p.C_Smag * o->p.C_Smag * o->delta * o->delta;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++)
            for (int k = 0; k < Mz; k++) {
                size_t ind = ((size_t)i * My + j) * 2 * nz + k;
                double ne = 2.0 * gu[0][ind] * gu[0][ind];
                ne += 2.0 * gv[1][ind] * gv[1][ind];
                ne += 2.0 * gw[2][ind] * gw[2][ind];
                ne += (gu[1][ind] + gv[0][ind]) * (gu[1][ind] + gv[0][ind]);
                ne += (gu[2][ind] + gw[0][ind]) * (gu[2][ind] + gw[0][ind]);
                ne += (gv[2][ind] + gw[1][ind]) * (gv[2][ind] + gw[1][ind]);
                ne = ct * sqrt(ne);
                o->sgs[0][ind] = (-2.0) * ne * gu[0][ind];
                o->sgs[1][ind] = (-1.0) * ne * (gu[1][ind] + gv[0][ind]);
                o->sgs[2][ind] = (-2.0) * ne * gv[1][ind];
                o->sgs[3][ind] = (-1.0) * ne * (gu[2][ind] + gw[0][ind]);
                o->sgs[4][ind] = (-1.0) * ne * (gv[2][ind] + gw[1][ind]);
                o->sgs[5][ind] = (-2.0) * ne * gw[2][ind];
            }
    /* six forward transforms; the divergence is accumulated in the
     * reference's operand order (src/HIT_LES.cpp:179-191) */
    double *sk[6];
    for (int s = 0; s < 6; s++) sk[s] = calloc_d(2 * o->nc);
    for (int s = 0; s < 6; s++) {
        r2c3d(&o->fft, o->sgs[s], sk[s]);
        normalize_D(o, sk[s]);
    }
    const double *S11 = sk[0], *S21 = sk[1], *S22 = sk[2], *S31 = sk[3], *S32 = sk[4], *S33 = sk[5];
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 <= o->Mz2; k3++) {
                size_t ind = CIX(o, b, a, k3);
                if (!o->dealiased[ind]) continue;
                int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
                size_t r = 2 * ind, im = 2 * ind + 1;
                o->div_k[0][r] = -k1 * S11[im] - k2 * S21[im] - k3 * S31[im];
                o->div_k[0][im] = k1 * S11[r] + k2 * S21[r] + k3 * S31[r];
                o->div_k[1][r] = -k1 * S21[im] - k2 * S22[im] - k3 * S32[im];
                o->div_k[1][im] = k1 * S21[r] + k2 * S22[r] + k3 * S32[r];
                o->div_k[2][r] = -k1 * S31[im] - k2 * S32[im] - k3 * S33[im];
                o->div_k[2][im] = k1 * S31[r] + k2 * S32[r] + k3 * S33[r];
            }
    for (int s = 0; s < 6; s++) free(sk[s]);
}

/* Shell integral (src/HIT_LES.cpp:1137-1214).  which: 0 kinetic energy
 * (src/HIT_LES.cpp:991-996), 1 enstrophy (:1000-1003).  If bins != NULL the
 * per-shell values are written for K in [Kmin,Kmax] and their sum returned;
 * otherwise the direct sum over modes is returned.  Sums accumulate in long
 * double: the oracle is the accuracy reference, and a plain sequential double
 * sum over 10^7..10^8 modes drifts by up to 1e-12 relative on its own
 * (summation order is not part of the reference's algorithm: its MPI
 * reduction order changes with the rank count). */
static double integrate_field(orc_t *o, int which, double *bins, int Kmin, int Kmax)
{
    const int Mx = o->p.Mx, My = o->p.My, nz = o->Mz2 + 1;
    const double *u = o->uk[0], *v = o->uk[1], *w = o->uk[2];
    if (bins) {
        long double *lb = (long double *)calloc((size_t)(Kmax + 1), sizeof(long double));
        for (int b = 0; b < My; b++)
            for (int a = 0; a < Mx; a++)
                for (int k3 = 0; k3 < nz; k3++) {
                    size_t ind = CIX(o, b, a, k3);
                    if (!o->dealiased[ind]) continue;
                    int K = shell_of((int)o->rad2[ind]);
                    if (K < Kmin || K > Kmax) continue;
                    size_t r = 2 * ind, im = r + 1;
                    double ke = 0.5 * (u[r] * u[r] + u[im] * u[im] + v[r] * v[r] + v[im] * v[im] + w[r] * w[r] + w[im] * w[im]);
                    double aux = (which == 0) ? ke : 2 * o->rad2[ind] * ke;
                    lb[K] += o->conjugate[ind] ? (2 * aux) : aux;
                }
        long double s = 0.0L;
        for (int K = Kmin; K <= Kmax; K++) {
            bins[K] = (double)lb[K];
            s += lb[K];
        }
        free(lb);
        return (double)s;
    }
    long double s = 0.0L;
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 < nz; k3++) {
                size_t ind = CIX(o, b, a, k3);
                if (!o->dealiased[ind]) continue;
                int K = shell_of((int)o->rad2[ind]);
                if (K < Kmin || K > Kmax) continue;
                size_t r = 2 * ind, im = r + 1;
                double ke = 0.5 * (u[r] * u[r] + u[im] * u[im] + v[r] * v[r] + v[im] * v[im] + w[r] * w[r] + w[im] * w[im]);
                double aux = (which == 0) ? ke : 2 * o->rad2[ind] * ke;
                s += o->conjugate[ind] ? (2 * aux) : aux;
            }
    return (double)s;
}

double orc_energy(orc_t *o) /* src/HIT_LES.cpp:1254-1259 */
{
    o->Ek_Tot = integrate_field(o, 0, NULL, 0, o->last_rad_max);
    return o->Ek_Tot;
}

double orc_enstrophy(orc_t *o) /* src/HIT_LES.cpp:490-495 */
{
    o->Ok_Tot = integrate_field(o, 1, NULL, 1, o->last_rad_max);
    return o->Ok_Tot;
}

double orc_energy_cascade(orc_t *o, double *Ek_out) /* src/HIT_LES.cpp:1261-1266 */
{
    o->Ek_Tot = integrate_field(o, 0, o->Ek, 0, o->last_rad_max);
    if (Ek_out) memcpy(Ek_out, o->Ek, sizeof(double) * (o->last_rad_max + 1));
    return o->Ek_Tot;
}

double orc_relambda(orc_t *o) /* src/HIT_LES.cpp:478-483 */
{
    orc_energy(o);
    orc_enstrophy(o);
    o->ReLambda = o->Ek_Tot / o->p.nu * sqrt(20.0 / 3.0 / o->Ok_Tot);
    return o->ReLambda;
}

/* R vector (src/HIT_LES.cpp:196-280) */
static void r_vector(orc_t *o)
{
    for (int c = 0; c < 3; c++) memcpy(o->R0[c], o->R1[c], sizeof(double) * 2 * o->nc);
    convective(o);
    const int Mx = o->p.Mx, My = o->p.My;
#pragma omp parallel for schedule(static)
    for (size_t ind = 0; ind < o->nc; ind++) {
        if (!o->dealiased[ind]) continue;
        double ct = o->p.nu * o->rad2[ind];
        int mean = !(o->rad2[ind] > 0.0);
        for (int ic = 0; ic <= 1; ic++) {
            size_t q = 2 * ind + ic;
            o->R1[0][q] = -(mean ? 0.0 : o->conv_k[0][q]) - ct * o->uk[0][q];
            o->R1[1][q] = -(mean ? 0.0 : o->conv_k[1][q]) - ct * o->uk[1][q];
            o->R1[2][q] = -(mean ? 0.0 : o->conv_k[2][q]) - ct * o->uk[2][q];
        }
    }
    if (o->p.isLES) {
        sgs_divergence(o);
#pragma omp parallel for schedule(static)
        for (size_t ind = 0; ind < o->nc; ind++) {
            if (!o->dealiased[ind]) continue;
            for (int ic = 0; ic <= 1; ic++) {
                size_t q = 2 * ind + ic;
                o->R1[0][q] -= o->div_k[0][q];
                o->R1[1][q] -= o->div_k[1][q];
                o->R1[2][q] -= o->div_k[2][q];
            }
        }
    }
    if (o->p.isRotating) {
        const double ox = o->p.omega_x, oy = o->p.omega_y, oz = o->p.omega_z;
#pragma omp parallel for schedule(static)
        for (size_t ind = 0; ind < o->nc; ind++) {
            if (!o->dealiased[ind]) continue;
            for (int ic = 0; ic <= 1; ic++) {
                size_t q = 2 * ind + ic;
                o->R1[0][q] += 2 * (oz * o->uk[1][q] - oy * o->uk[2][q]);
                o->R1[1][q] += 2 * (ox * o->uk[2][q] - oz * o->uk[0][q]);
                o->R1[2][q] += 2 * (oy * o->uk[0][q] - ox * o->uk[1][q]);
            }
        }
    }
    if (o->p.force_on) { /* src/HIT_LES.cpp:243-278 (disabled there by #if 0) */
        const int Kfmin = o->p.force_kmin, Kfmax = o->p.force_kmax;
        double Ef = integrate_field(o, 0, o->Ek, Kfmin, Kfmax);
        double alpha = 0.0;
        if (Ef > 1e-30) alpha = o->p.force_eps / (2.0 * Ef);
        (void)Mx; (void)My;
        for (size_t ind = 0; ind < o->nc; ind++) {
            int K = shell_of((int)o->rad2[ind]);
            if (!o->dealiased[ind]) continue;
            for (int ic = 0; ic <= 1; ic++) {
                size_t q = 2 * ind + ic;
                if (Kfmin <= K && K <= Kfmax) {
                    o->R1[0][q] += alpha * o->uk[0][q];
                    o->R1[1][q] += alpha * o->uk[1][q];
                    o->R1[2][q] += alpha * o->uk[2][q];
                } else if (K == 0) {
                    o->R1[0][q] = 0.0; o->R1[1][q] = 0.0; o->R1[2][q] = 0.0;
                }
            }
        }
    }
}

/* src/HIT_LES.cpp:295-339.  At the first self-adaptive step the reference
 * reads FFTW scratch for the oldest level; here it is the current field. */
static void predictor(orc_t *o)
{
    const double At = o->At;
    if (o->p.isSelfAdaptive) {
        for (int c = 0; c < 3; c++) {
            if (o->iter > 0) memcpy(o->ukaux[c], o->uk0[c], sizeof(double) * 2 * o->nc);
            else memcpy(o->ukaux[c], o->uk[c], sizeof(double) * 2 * o->nc);
            memcpy(o->uk0[c], o->uk[c], sizeof(double) * 2 * o->nc);
        }
        o->kappa05 = (1.0 / (o->kappa + 0.5));
        const double kap = o->kappa, k05 = o->kappa05;
#pragma omp parallel for schedule(static)
        for (size_t ind = 0; ind < o->nc; ind++) {
            if (!o->dealiased[ind]) continue;
            for (int ic = 0; ic <= 1; ic++) {
                size_t q = 2 * ind + ic;
                for (int c = 0; c < 3; c++)
                    o->uk[c][q] = k05 * (2.0 * kap * o->uk0[c][q] + At * ((1.0 + kap) * o->R1[c][q] - kap * o->R0[c][q])
                                          - (kap - 0.5) * o->ukaux[c][q]);
            }
        }
    } else {
#pragma omp parallel for schedule(static)
        for (size_t ind = 0; ind < o->nc; ind++) {
            if (!o->dealiased[ind]) continue;
            for (int ic = 0; ic <= 1; ic++) {
                size_t q = 2 * ind + ic;
                for (int c = 0; c < 3; c++)
                    o->uk[c][q] = o->uk[c][q] + At * (1.5 * o->R1[c][q] - 0.5 * o->R0[c][q]);
            }
        }
    }
}

/* src/HIT_LES.cpp:346-359 */
static void projection(orc_t *o)
{
    const int Mx = o->p.Mx, My = o->p.My;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 <= o->Mz2; k3++) {
                size_t ind = CIX(o, b, a, k3);
                if (!o->dealiased[ind]) continue;
                int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
                for (int ic = 0; ic <= 1; ic++) {
                    size_t q = 2 * ind + ic;
                    double aux = ((o->rad2[ind] > 0) ? ((k1 * o->uk[0][q] + k2 * o->uk[1][q] + k3 * o->uk[2][q]) / o->rad2[ind]) : 0.0);
                    o->uk[0][q] -= aux * k1;
                    o->uk[1][q] -= aux * k2;
                    o->uk[2][q] -= aux * k3;
                }
            }
}

/* self-adaptive helpers src/HIT_LES.h:249-269 */
static double FuL(double x, double x0, double x1, double f0, double f1) { return f0 + (x - x0) * (f1 - f0) / (x1 - x0); }
static double FuQ(double x, double x0, double x1) { return (x - x0) * (x - x1); }
static double FuG(double x, double a, double b, double c, double x0, double x1, double f0, double f1)
{
    return (a * x * x + b * x + c) * FuQ(x, x0, x1) + FuL(x, x0, x1, f0, f1);
}
static double Topt(orc_t *o, double p0)
{
    if (p0 >= 0.0 && p0 < o->phi[0]) return FuG(p0, 0.0, o->cc[0], o->cc[1], 0.0, o->phi[0], 4.0 / 3.0, o->t1);
    return FuG(p0, o->cc[2], o->cc[3], o->cc[4], o->phi[0], ORC_PI / 2.0, o->t1, 1.0);
}
static double Kopt(orc_t *o, double p0)
{
    if (p0 >= 0.0 && p0 <= o->phi[0]) return 1.0;
    if (p0 > o->phi[0] && p0 <= o->phi[1]) return FuG(p0, o->cc[5], o->cc[6], o->cc[7], o->phi[0], o->phi[1], 1.0, o->kk[0]);
    if (p0 > o->phi[1] && p0 <= o->phi[2]) return FuG(p0, o->cc[8], o->cc[9], o->cc[10], o->phi[1], o->phi[2], o->kk[0], o->kk[1]);
    return FuG(p0, o->cc[11], o->cc[12], o->cc[13], o->phi[2], ORC_PI / 2.0, o->kk[1], 0.0);
}

static double max_vel(orc_t *o, int sum_form)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    const double *u = o->u[0], *v = o->u[1], *w = o->u[2];
    double m = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : m)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++)
            for (int k = 0; k < Mz; k++) {
                size_t ind = ((size_t)i * My + j) * 2 * nz + k;
                double x;
                if (sum_form) x = fabs(u[ind]) / o->Ax + fabs(v[ind]) / o->Ay + fabs(w[ind]) / o->Az;
                else {
                    double b = fabs(v[ind]) / o->Ay, c = fabs(w[ind]) / o->Az;
                    double bc = b > c ? b : c;
                    double a = fabs(u[ind]) / o->Ax;
                    x = a > bc ? a : bc;
                }
                if (x > m) m = x;
            }
    return m;
}

/* src/HIT_LES.cpp:365-403; a fixed step replaces the CFL when set. */
void orc_recalc_timestep(orc_t *o)
{
    if (o->p.isSelfAdaptive) {
        o->MaxVel = max_vel(o, 1);
        const double nu = o->p.nu, Ah = o->Ah;
        o->phi0 = atan(o->MaxVel * Ah * Ah / (0.35 * nu * 12.0));
        o->At = Topt(o, o->phi0) / sqrt((o->MaxVel * o->MaxVel) / (0.35 * 0.35) + (144.0 * nu * nu / (Ah * Ah * Ah * Ah)));
        o->kappa = Kopt(o, o->phi0);
        o->At = o->p.C_At * o->At;
        return;
    }
    o->MaxVel = max_vel(o, 0);
    double ax2 = o->Ax * o->Ax, ay2 = o->Ay * o->Ay, az2 = o->Az * o->Az;
    double m = ay2 < az2 ? ay2 : az2;
    m = ax2 < m ? ax2 : m;
    o->C_visc = 0.2 * m / o->p.nu;
    o->C_conv = 0.35 / o->MaxVel;
    o->At = o->C_visc < o->C_conv ? o->C_visc : o->C_conv;
    o->At = o->p.C_At * o->At;
    if (o->p.fixed_dt > 0.0) o->At = o->p.fixed_dt;
}

/* src/HIT_LES.cpp:9-25 */
void orc_step(orc_t *o)
{
    if (o->uk_start[0])
        for (int c = 0; c < 3; c++) memcpy(o->uk_start[c], o->uk[c], sizeof(double) * 2 * o->nc);
    r_vector(o);
    predictor(o);
    projection(o);
    orc_recalc_timestep(o);
    o->time = o->time + o->At;
    o->iter++;
}

void orc_steps(orc_t *o, int n)
{
    for (int i = 0; i < n; i++) orc_step(o);
}

/* src/HIT_LES.cpp:485-489 */
void orc_forcing_relambda(orc_t *o, double ReLambda_)
{
    orc_energy(o);
    orc_enstrophy(o);
    o->p.nu = o->Ek_Tot / ReLambda_ * sqrt(20.0 / 3.0 / o->Ok_Tot);
}

/* src/HIT_LES.cpp:446-476 (one rank: the gather is the identity) */
void orc_ccc(orc_t *o)
{
    const int Mx = o->p.Mx, My = o->p.My;
    size_t plane = (size_t)Mx * My;
    double *copy = calloc_d(2 * plane * 3);
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int c = 0; c < 3; c++) {
                size_t ind = CIX(o, b, a, 0);
                copy[2 * (c * plane + (size_t)b * Mx + a)] = o->uk[c][2 * ind];
                copy[2 * (c * plane + (size_t)b * Mx + a) + 1] = o->uk[c][2 * ind + 1];
            }
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++) {
            int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
            if (k1 < 0 || (k1 == 0 && k2 < 0)) {
                int a_c = -k1, k2c = -k2;
                int b_c = (k2c >= 0) ? k2c : (My + k2c);
                size_t ind = CIX(o, b, a, 0);
                for (int c = 0; c < 3; c++) {
                    size_t src = c * plane + (size_t)b_c * Mx + a_c;
                    o->uk[c][2 * ind] = copy[2 * src];
                    o->uk[c][2 * ind + 1] = -copy[2 * src + 1];
                }
            }
        }
    free(copy);
}

/* src/HIT_LES.cpp:410-444 */
void orc_forcing_energy_cascade(orc_t *o, const double *Forced_Ek, long last_input_rad, int nullify)
{
    orc_energy_cascade(o, NULL);
    int last_eff = (int)(last_input_rad < o->last_rad ? last_input_rad : o->last_rad);
    double *FT = calloc_d((size_t)last_eff + 1);
    for (int K = 0; K <= last_eff; K++) FT[K] = ((o->Ek[K] > 0.0) ? sqrt(Forced_Ek[K] / o->Ek[K]) : 0.0);
    for (size_t ind = 0; ind < o->nc; ind++) {
        if (!o->dealiased[ind]) continue;
        int K = shell_of((int)o->rad2[ind]);
        for (int ic = 0; ic <= 1; ic++) {
            size_t q = 2 * ind + ic;
            if (K <= last_eff) {
                o->uk[0][q] *= FT[K]; o->uk[1][q] *= FT[K]; o->uk[2][q] *= FT[K];
            } else if (nullify) {
                o->uk[0][q] = 0.0; o->uk[1][q] = 0.0; o->uk[2][q] = 0.0;
            }
        }
    }
    free(FT);
    if (o->p.isCCC) orc_ccc(o);
}

/* K41 initial field (src/HIT_LES.cpp:502-540), emulating nranks FFTW slabs:
 * every rank calls srand(0) and draws glibc rand() over its own y-planes
 * (local_n1 = ceil(My/nranks), src/FFTW_package.cpp:49). */
void orc_init_k41(orc_t *o, int nranks)
{
    const int Mx = o->p.Mx, My = o->p.My;
    int block = (My + nranks - 1) / nranks;
    for (int r = 0; r < nranks; r++) {
        int b0 = r * block, b1 = b0 + block;
        if (b1 > My) b1 = My;
        srand(0);
        for (int b = b0; b < b1; b++)
            for (int a = 0; a < Mx; a++)
                for (int k3 = 0; k3 <= o->Mz2; k3++) {
                    size_t ind = CIX(o, b, a, k3);
                    for (int ic = 0; ic <= 1; ic++) {
                        size_t q = 2 * ind + ic;
                        if (o->dealiased[ind]) {
                            int K = shell_of((int)o->rad2[ind]);
                            if (K <= o->last_rad) {
                                o->uk[0][q] = (((double)rand() / RAND_MAX) * 2.0 - 1.0);
                                o->uk[1][q] = (((double)rand() / RAND_MAX) * 2.0 - 1.0);
                                o->uk[2][q] = (((double)rand() / RAND_MAX) * 2.0 - 1.0);
                            } else {
                                o->uk[0][q] = 0.0; o->uk[1][q] = 0.0; o->uk[2][q] = 0.0;
                            }
                        } else {
                            o->uk[0][q] = 0.0; o->uk[1][q] = 0.0; o->uk[2][q] = 0.0;
                        }
                    }
                }
    }
    orc_ccc(o);
    double *K41 = calloc_d((size_t)o->last_rad + 1);
    K41[0] = 0.0;
    for (int K = 1; K <= o->last_rad; K++) K41[K] = pow(K, -5.0 / 3.0);
    orc_forcing_energy_cascade(o, K41, o->last_rad, 0);
    free(K41);
}

/* Synthetic initial field (no reference counterpart; the seeded workload
 * input of the measured configurations).  Random-phase, solenoidal,
 * E(k) ~ k^4 exp(-2 (k/k0)^2), scaled to E_tot = 1.5 u_rms^2.
 *
 * Hash: SplitMix64's output function (Steele, Lea and Flood, "Fast
 * splittable pseudorandom number generators", OOPSLA 2014; the constants of
 * Vigna's splitmix64.c), applied to x + 0x9E3779B97F4A7C15.  Each mode's two
 * uniforms per component are keyed by (seed, k1, k2, k3, draw), so the field
 * does not depend on any loop or launch order.  Gaussians by Box-Muller;
 * the component along k is removed; each mode gets |u_hat|^2 =
 * E(|k|) / (4 pi |k|^2) before the global scaling; the k3 = 0 plane takes
 * the conjugate of the canonical half (k1 > 0, or k1 = 0 and k2 > 0). */
uint64_t orc_splitmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double orc_rand_unif(uint64_t seed, int k1, int k2, int k3, int d)
{
    uint64_t key = orc_splitmix(seed) ^ orc_splitmix(((uint64_t)(uint32_t)(k1 + 65536) << 40) ^
                                                     ((uint64_t)(uint32_t)(k2 + 65536) << 20) ^
                                                     (uint64_t)(uint32_t)k3 ^ ((uint64_t)d << 58));
    return ((double)(orc_splitmix(key) >> 11) + 0.5) * (1.0 / 9007199254740992.0); /* in (0,1) */
}

void orc_init_random(orc_t *o, uint64_t seed, double k0, double urms)
{
    const int Mx = o->p.Mx, My = o->p.My;
    for (int c = 0; c < 3; c++) memset(o->uk[c], 0, sizeof(double) * 2 * o->nc);
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 <= o->Mz2; k3++) {
                size_t ind = CIX(o, b, a, k3);
                if (!o->dealiased[ind]) continue;
                int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
                int canon = k3 > 0 || k1 > 0 || (k1 == 0 && k2 > 0);
                int mirror = (k3 == 0) && !canon && !(k1 == 0 && k2 == 0);
                if (!canon && !mirror) continue; /* k = 0 */
                if (mirror) { k1 = -k1; k2 = -k2; }
                const double kk[3] = {(double)k1, (double)k2, (double)k3};
                const double r2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
                double gr[3], gi[3];
                for (int c = 0; c < 3; c++) {
                    double u1 = orc_rand_unif(seed, k1, k2, k3, 2 * c), u2 = orc_rand_unif(seed, k1, k2, k3, 2 * c + 1);
                    double rr = sqrt(-2.0 * log(u1));
                    gr[c] = rr * cos(2 * ORC_PI * u2);
                    gi[c] = rr * sin(2 * ORC_PI * u2);
                }
                double dr = (kk[0] * gr[0] + kk[1] * gr[1] + kk[2] * gr[2]) / r2;
                double di = (kk[0] * gi[0] + kk[1] * gi[1] + kk[2] * gi[2]) / r2;
                double nrm = 0.0;
                for (int c = 0; c < 3; c++) {
                    gr[c] -= dr * kk[c];
                    gi[c] -= di * kk[c];
                    nrm += gr[c] * gr[c] + gi[c] * gi[c];
                }
                if (!(nrm > 0.0)) continue;
                double km = sqrt(r2);
                double Ek = km * km * km * km * exp(-2.0 * (km / k0) * (km / k0));
                double amp = sqrt(Ek / (4.0 * ORC_PI * r2) / nrm);
                for (int c = 0; c < 3; c++) {
                    o->uk[c][2 * ind] = gr[c] * amp;
                    o->uk[c][2 * ind + 1] = (mirror ? -1.0 : 1.0) * gi[c] * amp;
                }
            }
    double e = integrate_field(o, 0, NULL, 0, o->last_rad_max);
    double sc = e > 0 ? sqrt(1.5 * urms * urms / e) : 0.0;
    for (int c = 0; c < 3; c++) {
        for (size_t q = 0; q < 2 * o->nc; q++) o->uk[c][q] *= sc;
        memset(o->R1[c], 0, sizeof(double) * 2 * o->nc);
    }
    o->time = 0.0;
    o->iter = 0;
}

/* Physical file field -> Fourier coefficients (src/HIT_LES.cpp:542-599,
 * 738-763, cubic files).  u,v,w are [x][y][z] of size nf^3 each.  Also
 * returns the file spectrum's energy and enstrophy (src/HIT_LES.cpp:1268-1294). */
void orc_init_from_physical(orc_t *o, const double *uf, const double *vf, const double *wf,
                            int nxf, int nyf, int nzf, double *Ek_file, double *Ens_file)
{
    fft3d ff;
    fft3d_init(&ff, nxf, nyf, nzf);
    size_t nzh = nzf / 2 + 1;
    size_t nrf = (size_t)nxf * nyf * 2 * nzh, ncf = (size_t)nyf * nxf * nzh;
    double *rr = calloc_d(nrf), *cc[3];
    const double *src[3] = {uf, vf, wf};
    for (int c = 0; c < 3; c++) {
        cc[c] = calloc_d(2 * ncf);
        for (int i = 0; i < nxf; i++)
            for (int j = 0; j < nyf; j++)
                for (int k = 0; k < nzf; k++)
                    rr[((size_t)i * nyf + j) * 2 * nzh + k] = src[c][((size_t)i * nyf + j) * nzf + k];
        r2c3d(&ff, rr, cc[c]); /* layout [y][x][z] */
        double nf = (double)(nxf * nyf * nzf);
        for (size_t q = 0; q < 2 * ncf; q++) cc[c][q] /= nf; /* :577-583 (int divisor) */
    }
    /* energy of the file spectrum (:1268-1294) */
    double ekf = 0.0, ensf = 0.0;
    for (int a = 0; a < nxf; a++)
        for (int b = 0; b < nyf; b++)
            for (int k3 = 0; k3 < (int)nzh; k3++) {
                size_t id = ((size_t)b * nxf + a) * nzh + k3;
                int k1 = ffreq(a, nxf), k2 = ffreq(b, nyf);
                int r2 = k1 * k1 + k2 * k2 + k3 * k3;
                int rel = ((k1 % o->p.Lx == 0) && (k2 % o->p.Ly == 0) && (k3 % o->p.Lz == 0));
                if (!rel) continue;
                int conj = (k3 != 0) && !((nzf % 2 == 0) && (k3 == nzf / 2));
                double e = (cc[0][2 * id] * cc[0][2 * id]) + (cc[0][2 * id + 1] * cc[0][2 * id + 1]);
                e += (cc[1][2 * id] * cc[1][2 * id]) + (cc[1][2 * id + 1] * cc[1][2 * id + 1]);
                e += (cc[2][2 * id] * cc[2][2 * id]) + (cc[2][2 * id + 1] * cc[2][2 * id + 1]);
                e *= 0.5;
                if (conj) { ekf += 2 * e; ensf += 2 * (2 * r2 * e); }
                else { ekf += e; ensf += (2 * r2 * e); }
            }
    if (Ek_file) *Ek_file = ekf;
    if (Ens_file) *Ens_file = ensf;
    /* truncation (:738-763) */
    const int Mx = o->p.Mx, My = o->p.My;
    for (int b = 0; b < My; b++)
        for (int a = 0; a < Mx; a++)
            for (int k3 = 0; k3 <= o->Mz2; k3++) {
                size_t ind = CIX(o, b, a, k3);
                int k1 = ffreq(a, Mx), k2 = ffreq(b, My);
                for (int ic = 0; ic <= 1; ic++) {
                    size_t q = 2 * ind + ic;
                    double val[3] = {0.0, 0.0, 0.0};
                    if (o->dealiased[ind] && (abs(k1) <= (nxf - 1) / 2) && (abs(k2) <= (nyf - 1) / 2) && (k3 <= (nzf - 1) / 2)) {
                        int af = (k1 >= 0) ? k1 : (nxf + k1);
                        int bf = (k2 >= 0) ? k2 : (nyf + k2);
                        size_t idf = ((size_t)bf * nxf + af) * nzh + k3; /* our file spectrum is [y][x][z] */
                        for (int c = 0; c < 3; c++) val[c] = cc[c][2 * idf + ic];
                    }
                    for (int c = 0; c < 3; c++) o->uk[c][q] = val[c];
                }
            }
    for (int c = 0; c < 3; c++) free(cc[c]);
    free(rr);
    fft3d_free(&ff);
}

/* Completes HIT_init after the field is set (src/HIT_LES.cpp:1037-1086). */
void orc_finish_init(orc_t *o, int isReLambda)
{
    for (int c = 0; c < 3; c++) memset(o->R1[c], 0, sizeof(double) * 2 * o->nc);
    orc_antitransform(o);
    if (!isReLambda) orc_recalc_timestep(o);
    o->Ek_init = orc_energy(o);
}

/* ---------------------------- state exchange ---------------------------- */
/* Compact retained box in the reference's Fourier-dump order
 * (src/HIT_LES_output.cpp:46-53): b (k2) outer, a (k1), k3 inner, over
 * |k1|<=Nx/2, |k2|<=Ny/2, 0<=k3<=Nz/2 (the full box; modes masked by the
 * length factors are included and are zero). */
long orc_nmodes(const orc_t *o)
{
    return (long)(2 * o->Ny_2 + 1) * (2 * o->Nx_2 + 1) * (o->Nz_2 + 1);
}

static void compact_map(const orc_t *o, long *idx)
{
    const int Mx = o->p.Mx, My = o->p.My;
    long n = 0;
    for (int b = 0; b < My; b++) {
        int k2 = ffreq(b, My);
        if (abs(k2) > o->Ny_2) continue;
        for (int a = 0; a < Mx; a++) {
            int k1 = ffreq(a, Mx);
            if (abs(k1) > o->Nx_2) continue;
            for (int k3 = 0; k3 <= o->Nz_2; k3++) idx[n++] = (long)CIX(o, b, a, k3);
        }
    }
}

void orc_get_state(const orc_t *o, double *out /* 3 * nmodes * 2 */)
{
    long nm = orc_nmodes(o);
    long *idx = (long *)malloc(sizeof(long) * nm);
    compact_map(o, idx);
    for (int c = 0; c < 3; c++)
        for (long m = 0; m < nm; m++) {
            out[2 * (c * nm + m)] = o->uk[c][2 * idx[m]];
            out[2 * (c * nm + m) + 1] = o->uk[c][2 * idx[m] + 1];
        }
    free(idx);
}

void orc_set_state(orc_t *o, const double *in)
{
    long nm = orc_nmodes(o);
    long *idx = (long *)malloc(sizeof(long) * nm);
    compact_map(o, idx);
    for (int c = 0; c < 3; c++) {
        memset(o->uk[c], 0, sizeof(double) * 2 * o->nc);
        for (long m = 0; m < nm; m++) {
            if (!o->dealiased[idx[m]]) continue;
            o->uk[c][2 * idx[m]] = in[2 * (c * nm + m)];
            o->uk[c][2 * idx[m] + 1] = in[2 * (c * nm + m) + 1];
        }
    }
    free(idx);
}

/* R history (R1 of the last step) in the same compact order */
void orc_get_rhist(const orc_t *o, double *out)
{
    long nm = orc_nmodes(o);
    long *idx = (long *)malloc(sizeof(long) * nm);
    compact_map(o, idx);
    for (int c = 0; c < 3; c++)
        for (long m = 0; m < nm; m++) {
            out[2 * (c * nm + m)] = o->R1[c][2 * idx[m]];
            out[2 * (c * nm + m) + 1] = o->R1[c][2 * idx[m] + 1];
        }
    free(idx);
}

/* physical velocity (from the last antitransform) as [3][x][y][z] unpadded */
void orc_velocity_physical(const orc_t *o, double *out)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    size_t P = (size_t)Mx * My * Mz;
    for (int c = 0; c < 3; c++)
        for (int i = 0; i < Mx; i++)
            for (int j = 0; j < My; j++)
                for (int k = 0; k < Mz; k++)
                    out[c * P + ((size_t)i * My + j) * Mz + k] = o->u[c][((size_t)i * My + j) * 2 * nz + k];
}

/* QA Parseval self-check quantities (src/HIT_LES.cpp:1216-1252):
 * out[0]=EkFour (direct), out[1]=EkReal (physical), out[2]=EkSpNS */
void orc_parseval(orc_t *o, double *out)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    double ekf = 0.0;
    for (size_t ind = 0; ind < o->nc; ind++) {
        if (!o->dealiased[ind]) continue;
        double aux = 0.0;
        for (int c = 0; c < 3; c++) aux += o->uk[c][2 * ind] * o->uk[c][2 * ind] + o->uk[c][2 * ind + 1] * o->uk[c][2 * ind + 1];
        aux *= 0.5;
        ekf += o->conjugate[ind] ? 2 * aux : aux;
    }
    orc_antitransform(o);
    double ekr = 0.0;
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++)
            for (int k = 0; k < Mz; k++) {
                size_t ind = ((size_t)i * My + j) * 2 * nz + k;
                ekr += o->u[0][ind] * o->u[0][ind] + o->u[1][ind] * o->u[1][ind] + o->u[2][ind] * o->u[2][ind];
            }
    ekr = 0.5 * ekr / ((double)Mx * My * Mz);
    out[0] = ekf;
    out[1] = ekr;
    out[2] = orc_energy(o);
}

/* scalars: time, At, nu, MaxVel, Ek_Tot, Ok_Tot, ReLambda, skew_sum2, skew_sum3, iter, kappa */
void orc_get_scalars(const orc_t *o, double *s)
{
    s[0] = o->time; s[1] = o->At; s[2] = o->p.nu; s[3] = o->MaxVel;
    s[4] = o->Ek_Tot; s[5] = o->Ok_Tot; s[6] = o->ReLambda;
    s[7] = o->skew_sum2; s[8] = o->skew_sum3; s[9] = (double)o->iter; s[10] = o->kappa;
}

void orc_set_nu(orc_t *o, double nu) { o->p.nu = nu; }
int orc_last_rad_max(const orc_t *o) { return o->last_rad_max; }
int orc_last_rad(const orc_t *o) { return o->last_rad; }

/* ------------------------------------------------------------------------ */
/* Velocity-gradient invariant spectra (src/HIT_LES.cpp:1297-1319)           */
/* Invariants of G[i][j] = d u_i / d x_j (src/tensor_algebra.cpp:250-303):
 *   S = 0.5 G + 0.5 G^T (calc_force_symmetry), Omega = G - S (calc_Omega);
 *   Q(T) = 0.5 ((tr T)^2 - tr(T T)) (calc_second_invariant), R(T) = det T
 *   (calc_determinant, its term order);
 *   QG = Q(G), QS = Q(S), RG = det G, RS = det S,
 *   V2 = 4 (tr(S^2 Omega^2) - 2 QS QO), QO = QG - QS (calc_QO, :356).       */
static double tr3(double A[3][3]) { double t = 0.0; for (int i = 0; i <= 2; i++) t += A[i][i]; return t; }

static void mul3(double a[3][3], double b[3][3], double ab[3][3])
{
    for (int i = 0; i <= 2; i++)
        for (int j = 0; j <= 2; j++) {
            ab[i][j] = 0.0;
            for (int k = 0; k <= 2; k++) ab[i][j] += a[i][k] * b[k][j];
        }
}

static double second_inv(double T[3][3])
{
    double T2[3][3];
    mul3(T, T, T2);
    const double trT = tr3(T), trT2 = tr3(T2);
    return 0.5 * (trT * trT - trT2);
}

static double det3(double t[3][3])
{
    double det = 0.0;
    det += t[0][0] * t[1][1] * t[2][2];
    det += t[0][1] * t[1][2] * t[2][0];
    det += t[0][2] * t[1][0] * t[2][1];
    det -= t[0][2] * t[1][1] * t[2][0];
    det -= t[0][1] * t[1][0] * t[2][2];
    det -= t[0][0] * t[1][2] * t[2][1];
    return det;
}

/* g: G row-major (g[3 i + j] = d u_i / d x_j); out: QG QS RG RS V2 */
void orc_invariants_point(const double *g, double *out)
{
    double G[3][3], S[3][3], O[3][3], S2[3][3], O2[3][3], S2O2[3][3];
    for (int i = 0; i <= 2; i++)
        for (int j = 0; j <= 2; j++) G[i][j] = g[3 * i + j];
    for (int i = 0; i <= 2; i++)
        for (int j = 0; j <= 2; j++) S[i][j] = 0.5 * G[i][j] + 0.5 * G[j][i];
    for (int i = 0; i <= 2; i++)
        for (int j = 0; j <= 2; j++) O[i][j] = 1.0 * G[i][j] + -1.0 * S[i][j];
    const double QG = second_inv(G), QS = second_inv(S);
    mul3(S, S, S2);
    mul3(O, O, O2);
    mul3(S2, O2, S2O2);
    const double QO = QG - QS;
    out[0] = QG;
    out[1] = QS;
    out[2] = det3(G);
    out[3] = det3(S);
    out[4] = 4.0 * (tr3(S2O2) - 2.0 * QS * QO);
}

/* Keep the state at the start of every step (the field whose gradients the
 * reference's gradu/gradv/gradw hold after a step, src/HIT_LES.cpp:74-125). */
void orc_keep_start(orc_t *o)
{
    for (int c = 0; c < 3; c++)
        if (!o->uk_start[c]) {
            o->uk_start[c] = calloc_d(2 * o->nc);
            memcpy(o->uk_start[c], o->uk[c], sizeof(double) * 2 * o->nc);
        }
}

/* Shell spectra of the five invariants (src/HIT_LES.cpp:1297-1319 with the
 * VelGradInvariants integrands |I_k|^2, :1006-1025): the gradients of the
 * state at the start of the last step (the current state before any step
 * or without orc_keep_start), the invariants per point, r2c, normalisation
 * on D, and the weighted shell sums of Integrate_Field (:1137-1214) over
 * K = 0..last_rad_max, accumulated in long double.
 * out: 5 x (last_rad_max + 1), in the order QG QS RG RS V2. */
void orc_invariant_spectra(orc_t *o, double *out)
{
    const int Mx = o->p.Mx, My = o->p.My, Mz = o->p.Mz, nz = o->Mz2 + 1;
    const int nb = o->last_rad_max + 1;
    double *const *src = (o->iter > 0 && o->uk_start[0]) ? o->uk_start : o->uk;
    double *g[9], *inv[5];
    for (int f = 0; f < 9; f++) g[f] = calloc_d(o->nr);
    for (int f = 0; f < 5; f++) inv[f] = calloc_d(o->nr);
    double *ck = calloc_d(2 * o->nc);
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            gradient_spectrum(o, src[c], d, o->ctmp);
            c2r3d(&o->fft, o->ctmp, g[3 * c + d], ck);
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++)
            for (int k = 0; k < Mz; k++) {
                size_t ind = ((size_t)i * My + j) * 2 * nz + k;
                double gp[9], v[5];
                for (int f = 0; f < 9; f++) gp[f] = g[f][ind];
                orc_invariants_point(gp, v);
                for (int f = 0; f < 5; f++) inv[f][ind] = v[f];
            }
    for (int f = 0; f < 5; f++) {
        r2c3d(&o->fft, inv[f], ck);
        normalize_D(o, ck);
        long double *lb = (long double *)calloc((size_t)nb, sizeof(long double));
        for (size_t ind = 0; ind < o->nc; ind++) {
            if (!o->dealiased[ind]) continue;
            const int K = shell_of((int)o->rad2[ind]);
            const double aux = ck[2 * ind] * ck[2 * ind] + ck[2 * ind + 1] * ck[2 * ind + 1];
            lb[K] += o->conjugate[ind] ? (2 * aux) : aux;
        }
        for (int K = 0; K < nb; K++) out[(size_t)f * nb + K] = (double)lb[K];
        free(lb);
    }
    for (int f = 0; f < 9; f++) free(g[f]);
    for (int f = 0; f < 5; f++) free(inv[f]);
    free(ck);
}

/* --------------------------- FFT test hooks ----------------------------- */
int orc_fft_c2r(int Mx, int My, int Mz, const double *cin, double *rout)
{
    fft3d f;
    if (fft3d_init(&f, Mx, My, Mz)) return -1;
    size_t nc = (size_t)My * Mx * (Mz / 2 + 1);
    double *w = calloc_d(2 * nc);
    c2r3d(&f, cin, rout, w);
    free(w);
    fft3d_free(&f);
    return 0;
}

int orc_fft_r2c(int Mx, int My, int Mz, const double *rin, double *cout)
{
    fft3d f;
    if (fft3d_init(&f, Mx, My, Mz)) return -1;
    r2c3d(&f, rin, cout);
    fft3d_free(&f);
    return 0;
}

int orc_fft1d(int n, double *x, int sign)
{
    plan1d p;
    if (plan1d_init(&p, n)) return -1;
    double *w = calloc_d(2 * (size_t)n);
    fft1d(&p, x, w, sign);
    free(w);
    plan1d_free(&p);
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
