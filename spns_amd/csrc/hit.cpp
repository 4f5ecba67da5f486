// C++ facade (reference class surface) over include/spns.h.  File formats
// follow src/HIT_LES.cpp:604-732 (input) and src/HIT_LES_output.cpp:20-115
// (output); the spectrum file follows src/HIT_LES.cpp:1180-1187.
#include "hit.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace spns_hit {

// ------------------------------------------------------------------------
FFT_package::FFT_package(const ptrdiff_t Mx_, const ptrdiff_t My_, const ptrdiff_t Mz_, int precision, int device)
{
    int rc = spns_fft_create((int32_t)Mx_, (int32_t)My_, (int32_t)Mz_, precision, device, &f);
    if (rc) throw Error(rc, std::string("FFT_package: ") + spns_fft_last_error(nullptr));
    alloc_local = My_ * Mx_ * (Mz_ / 2 + 1);
    local_n0 = Mx_;
    local_n1 = My_;
}

FFT_package::~FFT_package()
{
    if (f) spns_fft_destroy(f);
}

void FFT_package::FFTWr2c(double* rin, double* cout)
{
    int rc = spns_fft_r2c(f, rin, cout);
    if (rc) throw Error(rc, spns_fft_last_error(f));
}

void FFT_package::IFFTWc2r(double* cin, double* rout)
{
    int rc = spns_fft_c2r(f, cin, rout);
    if (rc) throw Error(rc, spns_fft_last_error(f));
}

void FFT_package::Fourier_Normalization(double* phi, ptrdiff_t Nx, ptrdiff_t Ny, ptrdiff_t Nz)
{
    int rc = spns_fft_normalize(f, phi, (int32_t)Nx, (int32_t)Ny, (int32_t)Nz);
    if (rc) throw Error(rc, spns_fft_last_error(f));
}

void FFT_package::synchronize()
{
    int rc = spns_fft_synchronize(f);
    if (rc) throw Error(rc, spns_fft_last_error(f));
}

// ------------------------------------------------------------------------
static int actual(ptrdiff_t n, int L) { return (int)((n % 2 == 0) ? n / L : (n - 1) / L + 1); }  // src/HIT_LES.h:31-36

HIT::HIT(const ptrdiff_t& Nx_, const ptrdiff_t& Ny_, const ptrdiff_t& Nz_, const ptrdiff_t& Mx_, const ptrdiff_t& My_,
         const ptrdiff_t& Mz_, const double& nu_, const double& C_Smag_, const double& C_At_, const int& Lx_factor_,
         const int& Ly_factor_, const int& Lz_factor_, const double& omega_x_, const double& omega_y_,
         const double& omega_z_, const char* ASCII_Input_Filename_, const char* Binary_Input_Filename_,
         const ptrdiff_t& Nx_file_, const ptrdiff_t& Ny_file_, const ptrdiff_t& Nz_file_, const bool& isReLambda_,
         const bool& isInitialFieldFromBinaryFile_, const bool& isInitialFieldFromASCIIFile_, const bool& isNotCubic_,
         const bool& isRotating_, const bool& isLES_, const bool& isComplexConjugateCorrected_,
         const bool& isSelfAdaptive_, const Options& opt)
    : Nx(Nx_), Ny(Ny_), Nz(Nz_), Mx(Mx_), My(My_), Mz(Mz_), Nx_file(Nx_file_), Ny_file(Ny_file_), Nz_file(Nz_file_),
      Lx(Lx_factor_), Ly(Ly_factor_), Lz(Lz_factor_), isNotCubic(isNotCubic_),
      isBinary(isInitialFieldFromBinaryFile_), isASCII(isInitialFieldFromASCIIFile_), nu(nu_)
{
    spns_config c{};
    c.Nx = (int32_t)Nx; c.Ny = (int32_t)Ny; c.Nz = (int32_t)Nz;
    c.Mx = (int32_t)Mx; c.My = (int32_t)My; c.Mz = (int32_t)Mz;
    c.nu = nu_;
    c.C_Smag = isLES_ ? C_Smag_ : 0.0;
    c.C_At = C_At_;
    c.Lx = Lx; c.Ly = Ly; c.Lz = Lz;
    c.omega_x = isRotating_ ? omega_x_ : 0.0;
    c.omega_y = isRotating_ ? omega_y_ : 0.0;
    c.omega_z = isRotating_ ? omega_z_ : 0.0;
    c.isNotCubic = isNotCubic_;
    c.isSelfAdaptive = isSelfAdaptive_;
    c.isComplexConjugateCorrected = isComplexConjugateCorrected_;
    c.precision = opt.precision;
    c.delta = opt.delta;
    c.fixed_dt = opt.fixed_dt;
    c.force_on = opt.force_on;
    c.force_eps = opt.force_eps;
    c.force_kmin = opt.force_kmin;
    c.force_kmax = opt.force_kmax;
    c.device = opt.device;
    c.use_graph = opt.use_graph;
    c.gradient_invariants = opt.gradient_invariants;
    c.keep_step_start = opt.keep_step_start;
    int rc = spns_create(&c, &ctx);
    if (rc) throw Error(rc, std::string("HIT: ") + spns_last_error(nullptr));
    ActualMx = actual(Mx, Lx); ActualMy = actual(My, Ly); ActualMz = actual(Mz, Lz);
    last_rad = spns_last_rad(ctx);
    Ek.assign(spns_nbins(ctx), 0.0);
    if (ASCII_Input_Filename_) ascii_name = ASCII_Input_Filename_;
    if (Binary_Input_Filename_) binary_name = Binary_Input_Filename_;
    // initial field (src/HIT_LES.cpp:1030-1035)
    if (isBinary || isASCII) {
        ActualNx_file = actual(Nx_file, Lx); ActualNy_file = actual(Ny_file, Ly); ActualNz_file = actual(Nz_file, Lz);
        std::vector<double> u, v, w;
        read_input_file(u, v, w);
        check(spns_init_from_physical(ctx, u.data(), v.data(), w.data(), (int32_t)Nx_file, (int32_t)Ny_file,
                                      (int32_t)Nz_file, &Ek_init_file, &Enstrophy_init_file),
              "Input_Real_Field");
    } else if (opt.random_init) {
        check(spns_init_random(ctx, opt.seed, opt.k0, opt.urms), "init_random");
    } else {
        check(spns_init_k41(ctx, opt.k41_ranks), "Input_K41_Field");
    }
    // rest of HIT_init (src/HIT_LES.cpp:1037-1086)
    check(spns_finish_init(ctx, isReLambda_), "HIT_init");
    Ek_init = Recalculate_Energy();
    qa_check = opt.qa_check;
}

HIT::~HIT()
{
    if (ctx) spns_destroy(ctx);
}

void HIT::check(int rc, const char* what) const
{
    if (rc) throw Error(rc, std::string(what) + ": " + spns_last_error(ctx));
}

void HIT::refresh() const
{
    spns_scalars s;
    check(spns_get_scalars(ctx, &s), "get_scalars");
    At = s.dt;
    time = s.time;
}

// Read_Input_File_Vxyz (src/HIT_LES.cpp:604-732): binary files hold u, v, w
// blocks of float or double (told apart by the file size); ASCII files hold
// rows "u v w".  Non-cubic runs read the Actual*_file box and repeat it.
void HIT::read_input_file(std::vector<double>& u, std::vector<double>& v, std::vector<double>& w) const
{
    const size_t nread = (size_t)ActualNx_file * ActualNy_file * ActualNz_file;
    std::vector<double> ur(nread, 0.0), vr(nread, 0.0), wr(nread, 0.0);
    if (isBinary) {
        std::ifstream in(binary_name, std::ios::binary);
        if (!in) throw Error(SPNS_EINVAL, "Error opening \"" + binary_name + "\"");
        in.seekg(0, std::ios::end);
        const size_t size = (size_t)in.tellg();
        in.seekg(0, std::ios::beg);
        if (size == sizeof(double) * 3 * nread) {
            in.read((char*)ur.data(), sizeof(double) * nread);
            in.read((char*)vr.data(), sizeof(double) * nread);
            in.read((char*)wr.data(), sizeof(double) * nread);
        } else if (size == sizeof(float) * 3 * nread) {
            std::vector<float> f(nread);
            for (std::vector<double>* dst : {&ur, &vr, &wr}) {
                in.read((char*)f.data(), sizeof(float) * nread);
                for (size_t i = 0; i < nread; i++) (*dst)[i] = f[i];
            }
        } else {
            throw Error(SPNS_EINVAL, "Incorrect size of \"" + binary_name + "\": " + std::to_string(size) +
                                         " bytes instead of " + std::to_string(sizeof(float) * 3 * nread) + " or " +
                                         std::to_string(sizeof(double) * 3 * nread));
        }
    } else {
        std::ifstream in(ascii_name);
        if (!in) throw Error(SPNS_EINVAL, "Error opening \"" + ascii_name + "\"");
        size_t row = 0;
        while (row < nread && (in >> ur[row] >> vr[row] >> wr[row])) row++;
    }
    const size_t nf = (size_t)Nx_file * Ny_file * Nz_file;
    u.assign(nf, 0.0); v.assign(nf, 0.0); w.assign(nf, 0.0);
    if (!isNotCubic) {
        u = ur; v = vr; w = wr;
        u.resize(nf); v.resize(nf); w.resize(nf);
        return;
    }
    for (ptrdiff_t i = 0; i < Nx_file; i++)
        for (ptrdiff_t j = 0; j < Ny_file; j++)
            for (ptrdiff_t k = 0; k < Nz_file; k++) {
                const size_t ind = ((size_t)i * Ny_file + j) * Nz_file + k;
                if (i < ActualNx_file && j < ActualNy_file && k < ActualNz_file) {
                    const size_t ir = ((size_t)i * ActualNy_file + j) * ActualNz_file + k;
                    u[ind] = ur[ir]; v[ind] = vr[ir]; w[ind] = wr[ir];
                } else {
                    const ptrdiff_t i_ = (i < ActualNx_file) ? i : i - ActualNx_file;
                    const ptrdiff_t j_ = (j < ActualNy_file) ? j : j - ActualNy_file;
                    const ptrdiff_t k_ = (k < ActualNz_file) ? k : k - ActualNz_file;
                    const size_t s = ((size_t)i_ * Ny_file + j_) * Nz_file + k_;
                    u[ind] = u[s]; v[ind] = v[s]; w[ind] = w[s];
                }
            }
}

void HIT::Check_PhysFour_Energy(const double atol)
{
    check(spns_check_phys_four_energy(ctx, atol, nullptr, nullptr, nullptr), "Check_PhysFour_Energy");
}

void HIT::New_Fractional_Step_Method()
{
    if (qa_check) Check_PhysFour_Energy(1e-13);  // src/HIT_LES.cpp:10-12 (QA build)
    check(spns_step(ctx, 1), "New_Fractional_Step_Method");
    phys_current = false;  // the step's u,v,w are those of its start state
}
void HIT::Recalculate_TimeStep() { check(spns_recalculate_timestep(ctx), "Recalculate_TimeStep"); }

void HIT::Forcing_Energy_Cascade(double const* const Forced_Ek, const ptrdiff_t& last_input_rad,
                                 const bool& isNullifyMissingEk)
{
    check(spns_forcing_energy_cascade(ctx, Forced_Ek, last_input_rad, isNullifyMissingEk), "Forcing_Energy_Cascade");
}

void HIT::Complex_Conjugate_Correction() { check(spns_complex_conjugate_correction(ctx), "Complex_Conjugate_Correction"); }

double HIT::Recalculate_Energy()
{
    spns_scalars s;
    check(spns_get_scalars(ctx, &s), "Recalculate_Energy");
    Ek_Tot = s.ek;
    return Ek_Tot;
}

double HIT::Recalculate_Enstrophy()
{
    spns_scalars s;
    check(spns_get_scalars(ctx, &s), "Recalculate_Enstrophy");
    Ok_Tot = s.om;
    return Ok_Tot;
}

double HIT::Recalculate_Reynolds_Lambda()
{
    Recalculate_Energy();
    Recalculate_Enstrophy();
    ReLambda = Ek_Tot / nu * std::sqrt(20.0 / 3.0 / Ok_Tot);
    return ReLambda;
}

void HIT::Forcing_Reynolds_Lambda(const double ReLambda_)
{
    Recalculate_Energy();
    Recalculate_Enstrophy();
    nu = Ek_Tot / ReLambda_ * std::sqrt(20.0 / 3.0 / Ok_Tot);
    check(spns_set_nu(ctx, nu), "Forcing_Reynolds_Lambda");
}

void HIT::setnu(const double& nu_)
{
    nu = nu_;
    check(spns_set_nu(ctx, nu), "setnu");
}

double HIT::Recalculate_Energy_Cascade(const char* filename)
{
    check(spns_energy_cascade(ctx, Ek.data(), (int32_t)Ek.size(), &Ek_Tot), "Recalculate_Energy_Cascade");
    if (filename) {
        std::ofstream out(filename);
        if (!out.is_open()) std::printf("WARNING: HIT::Integrate_Field(): Error opening \"%s\".\n", filename);
        for (size_t K = 0; K < Ek.size(); K++) out << K << "\t" << Ek[K] << "\n";
    }
    return Ek_Tot;
}

// Shell spectra of the five velocity-gradient invariants, one "K\tvalue" file
// each (src/HIT_LES.cpp:1137-1190 output format), of the gradients at the
// start of the last step; needs Options::gradient_invariants.
void HIT::Recalculate_Invariants_Distribution(const char* filename[5])
{
    const size_t nb = Ek.size();
    VelGradInv.assign(5 * nb, 0.0);
    check(spns_invariant_spectra(ctx, (int32_t)nb, VelGradInv.data()), "Recalculate_Invariants_Distribution");
    if (!filename) return;
    for (int f = 0; f < 5; f++) {
        if (!filename[f]) continue;
        std::ofstream out(filename[f]);
        if (!out.is_open()) std::printf("WARNING: HIT::Integrate_Field(): Error opening \"%s\".\n", filename[f]);
        for (size_t K = 0; K < nb; K++) out << K << "\t" << VelGradInv[f * nb + K] << "\n";
    }
}

void HIT::Recalculate_Velocity_Antitransform()
{
    // the CFL maximum of the next Recalculate_TimeStep then refers to the
    // current state's physical field, as in the reference
    check(spns_recalculate_velocity_antitransform(ctx), "Recalculate_Velocity_Antitransform");
    phys_current = true;
}

double HIT::getSkewness() const
{
    spns_scalars s;
    check(spns_get_scalars(ctx, &s), "getSkewness");
    return s.skewness;
}

// DealiasedComplex3DimField_to_BinaryFile (src/HIT_LES_output.cpp:20-63):
// for each component, the dealiased modes in LOOP_FOURIER order as (re,im)
// doubles; each component block holds all of its modes.
void HIT::FourierVelocity_to_BinaryFile(char const* const filename) const
{
    const int64_t nm = spns_nmodes(ctx);
    std::vector<double> st(6 * (size_t)nm);
    check(spns_get_state(ctx, st.data()), "FourierVelocity_to_BinaryFile");
    std::ofstream out(filename, std::ios::binary | std::ios::trunc);
    if (!out) throw Error(SPNS_EINVAL, std::string("cannot open ") + filename);
    const int nk1 = 2 * (int)(Nx / 2) + 1, nk2 = 2 * (int)(Ny / 2) + 1, nk3 = (int)(Nz / 2) + 1;
    for (int c = 0; c < 3; c++)
        for (int i2 = 0; i2 < nk2; i2++)
            for (int i1 = 0; i1 < nk1; i1++)
                for (int i3 = 0; i3 < nk3; i3++) {
                    const int k1 = i1 <= nk1 / 2 ? i1 : i1 - nk1, k2 = i2 <= nk2 / 2 ? i2 : i2 - nk2;
                    if (isNotCubic && ((k1 % Lx) || (k2 % Ly) || (i3 % Lz))) continue;  // not dealiased
                    const size_t q = (size_t)c * nm + ((size_t)i2 * nk1 + i1) * nk3 + i3;
                    out.write((const char*)&st[2 * q], 2 * sizeof(double));
                }
}

// Real3DimField_to_BinaryFile (src/HIT_LES_output.cpp:66-115): u, v, w
// blocks, each [x][y][z]; non-cubic runs keep x<ActualMx, y<ActualMy, z<ActualMz.
// The field is what the reference's u,v,w arrays hold: of the current state
// before any step and after Recalculate_Velocity_Antitransform, else of the
// last step's start state (needs Options::keep_step_start).
void HIT::RealVelocity_to_BinaryFile(char const* const filename) const
{
    const size_t P = (size_t)Mx * My * Mz;
    std::vector<double> uvw(3 * P);
    if (phys_current)
        check(spns_velocity_physical(ctx, uvw.data()), "RealVelocity_to_BinaryFile");
    else
        check(spns_velocity_physical_step_start(ctx, uvw.data()), "RealVelocity_to_BinaryFile");
    std::ofstream out(filename, std::ios::binary | std::ios::trunc);
    if (!out) throw Error(SPNS_EINVAL, std::string("cannot open ") + filename);
    for (int c = 0; c < 3; c++)
        for (ptrdiff_t i = 0; i < Mx; i++)
            for (ptrdiff_t j = 0; j < My; j++)
                for (ptrdiff_t k = 0; k < Mz; k++) {
                    if (isNotCubic && !(i < ActualMx && j < ActualMy && k < ActualMz)) continue;
                    out.write((const char*)&uvw[c * P + ((size_t)i * My + j) * Mz + k], sizeof(double));
                }
}

}  // namespace spns_hit
