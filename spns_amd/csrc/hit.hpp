// C++ facade with the reference's class surface (src/HIT_LES.h:21-246,
// src/FFTW_package.h:21-71) over the C ABI of include/spns.h, so a caller of
// SpNS's HIT / FFT_package keeps its code.  One GPU, one rank: the
// "local" MPI sizes are the whole grid.  Errors throw spns_hit::Error (the
// reference calls crash(), which prints and exits).
#pragma once
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spns.h"

namespace spns_hit {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// FFT_package (src/FFTW_package.h:21-71) on device buffers.
class FFT_package {
public:
    FFT_package(const ptrdiff_t Mx_, const ptrdiff_t My_, const ptrdiff_t Mz_, int precision = SPNS_PREC_FP64,
                int device = 0);
    ~FFT_package();
    FFT_package(const FFT_package&) = delete;
    FFT_package& operator=(const FFT_package&) = delete;
    // src/FFTW_package.cpp:7-14 (inputs destroyed, device pointers)
    void FFTWr2c(double* rin, double* cout);
    void IFFTWc2r(double* cin, double* rout);
    // src/FFTW_package.cpp:33-42 with the retained box of (Nx,Ny,Nz) as "relevant"
    void Fourier_Normalization(double* phi, ptrdiff_t Nx, ptrdiff_t Ny, ptrdiff_t Nz);
    void synchronize();
    const ptrdiff_t& getalloc_local() const { return alloc_local; }
    const ptrdiff_t& getlocal_n0() const { return local_n0; }
    const ptrdiff_t& getlocal_0_start() const { return zero; }
    const ptrdiff_t& getlocal_n1() const { return local_n1; }
    const ptrdiff_t& getlocal_1_start() const { return zero; }

private:
    spns_fft* f = nullptr;
    ptrdiff_t alloc_local, local_n0, local_n1, zero = 0;
};

// Build options beyond the reference's constructor (all optional).
struct Options {
    int precision = SPNS_PREC_FP64;
    double delta = 0.0;       // LES filter width; <= 0 reference formula
    double fixed_dt = 0.0;    // > 0 fixed step
    bool force_on = false;    // constant-power forcing
    double force_eps = 0.1;
    int force_kmin = 1, force_kmax = 2;
    bool random_init = false; // synthetic k^4 exp(-2(k/k0)^2) field instead of K41
    unsigned long long seed = 0;
    double k0 = 4.0, urms = 1.0;
    int k41_ranks = 1;        // emulate the per-rank K41 draws of an n-rank run
    int device = 0;
    bool use_graph = true;
    bool gradient_invariants = false;  // enables Recalculate_Invariants_Distribution
    bool keep_step_start = false;      // RealVelocity_to_BinaryFile between steps (the step-start field)
    bool qa_check = false;             // the QA build's Check_PhysFour_Energy(1e-13) before every step
};

// HIT (src/HIT_LES.h:21-246).
class HIT {
public:
    HIT(const ptrdiff_t& Nx_, const ptrdiff_t& Ny_, const ptrdiff_t& Nz_, const ptrdiff_t& Mx_, const ptrdiff_t& My_,
        const ptrdiff_t& Mz_, const double& nu_, const double& C_Smag_, const double& C_At_, const int& Lx_factor_,
        const int& Ly_factor_, const int& Lz_factor_, const double& omega_x_, const double& omega_y_, const double& omega_z_,
        const char* ASCII_Input_Filename_, const char* Binary_Input_Filename_, const ptrdiff_t& Nx_file_,
        const ptrdiff_t& Ny_file_, const ptrdiff_t& Nz_file_, const bool& isReLambda_,
        const bool& isInitialFieldFromBinaryFile_, const bool& isInitialFieldFromASCIIFile_, const bool& isNotCubic_,
        const bool& isRotating_, const bool& isLES_, const bool& isComplexConjugateCorrected_, const bool& isSelfAdaptive_,
        const Options& opt = Options());
    ~HIT();
    HIT(const HIT&) = delete;
    HIT& operator=(const HIT&) = delete;

    void New_Fractional_Step_Method();                      // src/HIT_LES.cpp:9-25
    void Recalculate_TimeStep();                            // :365-371
    void Forcing_Energy_Cascade(double const* const Forced_Ek, const ptrdiff_t& last_input_rad,
                                const bool& isNullifyMissingEk);  // :410-444
    void Complex_Conjugate_Correction();                    // :446-476
    double Recalculate_Reynolds_Lambda();                   // :478-483
    void Forcing_Reynolds_Lambda(const double ReLambda_);   // :485-489
    double Recalculate_Enstrophy();                         // :490-495
    double Recalculate_Energy();                            // :1254-1259
    double Recalculate_Energy_Cascade(const char* filename = nullptr);  // :1261-1266
    void Recalculate_Velocity_Antitransform();              // :60-69
    void Check_PhysFour_Energy(const double atol);          // :1216-1252 (QA)
    void Recalculate_Invariants_Distribution(const char* filename[5] = nullptr);  // :1297-1319
    void FourierVelocity_to_BinaryFile(char const* const filename) const;  // src/HIT_LES_output.cpp:8-10
    void RealVelocity_to_BinaryFile(char const* const filename) const;     // src/HIT_LES_output.cpp:12-14

    const ptrdiff_t& getNx() const { return Nx; }
    const ptrdiff_t& getNy() const { return Ny; }
    const ptrdiff_t& getNz() const { return Nz; }
    const ptrdiff_t& getlast_rad() const { return last_rad; }
    const double& getnu() const { return nu; }
    void setnu(const double& nu_);
    const double& getReLambda() const { return ReLambda; }
    const double& getAt() const { refresh(); return At; }
    const double& gettime() const { refresh(); return time; }
    const double& getOk_Tot() const { return Ok_Tot; }
    const double& getEk_Tot() const { return Ek_Tot; }
    const double& getEk_init() const { return Ek_init; }
    const double& getEk_init_file() const { return Ek_init_file; }
    const double& getEnstrophy_init_file() const { return Enstrophy_init_file; }
    const double& getEk(ptrdiff_t K) const { return Ek[K]; }
    double getSkewness() const;
    // shell spectrum K of invariant f (QG, QS, RG, RS, V2) from the last
    // Recalculate_Invariants_Distribution
    double getVelGradInv(int f, ptrdiff_t K) const { return VelGradInv[(size_t)f * Ek.size() + K]; }
    spns_ctx* context() const { return ctx; }

private:
    void check(int rc, const char* what) const;
    void refresh() const;  // pulls time/At from the device
    void read_input_file(std::vector<double>& u, std::vector<double>& v, std::vector<double>& w) const;

    spns_ctx* ctx = nullptr;
    ptrdiff_t Nx, Ny, Nz, Mx, My, Mz, last_rad, Nx_file, Ny_file, Nz_file;
    int Lx, Ly, Lz, ActualMx, ActualMy, ActualMz, ActualNx_file, ActualNy_file, ActualNz_file;
    bool isNotCubic, isBinary, isASCII;
    std::string ascii_name, binary_name;
    double nu, ReLambda = 0.0, Ok_Tot = 0.0, Ek_Tot = 0.0, Ek_init = 0.0, Ek_init_file = 0.0, Enstrophy_init_file = 0.0;
    mutable double At = 0.0, time = 0.0;
    bool phys_current = true;  // u,v,w of the current state (else of the last step's start)
    bool qa_check = false;     // Options::qa_check
    std::vector<double> Ek, VelGradInv;
};

}  // namespace spns_hit
