// spns_run: the MAIN.cpp-equivalent driver on one MI355X.
//
//   spns_run [parameters-file] [--qa]
//
// Reads the reference's parameters file (src/MAIN.cpp:127-172), derives the
// run flags and grid (src/MAIN.cpp:174-318), builds the HIT facade, sets nu
// from Re_lambda if asked (src/MAIN.cpp:336-357) and runs the main loop in the
// reference's order (src/MAIN.cpp:413-485): step, first-iteration Re_lambda
// forcing, spectrum forcing, periodic Re_lambda forcing, iteration count,
// logs, outputs.  With --qa (or QA 1 in the file) every step prints the QA
// log row of src/MAIN.cpp:440 and, as the reference's QA build does, first
// checks that the physical, spectral and shell-sum energies agree to 1e-13
// (HIT::Check_PhysFour_Energy); a violation ends the run with status 1.
// Output files use the reference's names.
//
// Optional keys beyond the reference: Precision (fp64 | fp32split),
// Fixed_Timestep, Max_Iterations, Delta, Constant_Power (injection rate),
// Forcing_Kmin, Forcing_Kmax, Random_Seed + Random_k0 (synthetic initial
// field), K41_Ranks (emulate the K41 draws of an n-rank run).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "hit.hpp"
#include "params.hpp"

using namespace spns_hit;

static int lcm2(int a, int b) { return a / std::gcd(a, b) * b; }

static std::vector<double> read_ek(const std::string& name, int last_K)  // src/MAIN.cpp:564-592
{
    std::ifstream in(name);
    if (!in) throw std::runtime_error("Read_Ek_Input_File(): Forced energy cascade input file is missing!!! Include \"" + name + "\", please");
    std::vector<double> ek(last_K + 1, 0.0);
    int K;
    double val;
    while (in >> K >> val) {
        if (K < 0 || K > last_K) throw std::runtime_error("Read_Ek_Input_File(): K out of range in \"" + name + "\"");
        ek[K] = val;
    }
    if (ek[0] > 0.0) throw std::runtime_error("Read_Ek_Input_File(): Ek_input[0] > 0 is not allowed");
    return ek;
}

int main(int argc, char** argv)
{
    std::string fparam = "./parameters";
    bool qa = false;
    for (int i = 1; i < argc; i++) {
        if (std::strcmp(argv[i], "--qa") == 0) qa = true;
        else fparam = argv[i];
    }
    try {
        Params PM(fparam);
        if (PM.integer("QA", 0)) qa = true;
        long Nx = PM.need_int("Nx"), Ny = PM.need_int("Ny"), Nz = PM.need_int("Nz");
        const double Final_Time = PM.need_real("Final_Time");
        if (PM.has("nu") && PM.has("Reynolds_lambda"))
            throw std::runtime_error("Kinematic viscosity and Reynolds lambda cannot be defined simoultaniously!!!");
        if (!PM.has("nu") && !PM.has("Reynolds_lambda"))
            throw std::runtime_error("Either kinematic viscosity or Reynolds lambda has to be defined!!!");
        double nu = PM.real("nu", 0.0);
        const double ReLambda = PM.real("Reynolds_lambda", 0.0);
        const long ReLambda_Freq = PM.integer("Reynolds_lambda_Freq", 0);
        long Mx = (3 * Nx) / 2, My = (3 * Ny) / 2, Mz = (3 * Nz) / 2;
        const bool isReLambda = PM.has("Reynolds_lambda");
        const bool isReLambda_iter = isReLambda && PM.has("Reynolds_lambda_Freq") && ReLambda_Freq > 0;
        long Nx_file = -1, Ny_file = -1, Nz_file = -1;
        bool isBin = false, isAsc = false;
        if (PM.has("Nx_file") && PM.has("Ny_file") && PM.has("Nz_file")) {
            Nx_file = PM.integer("Nx_file"); Ny_file = PM.integer("Ny_file"); Nz_file = PM.integer("Nz_file");
            if (PM.has("Binary_Input_Filename")) {
                isBin = true;
                if (PM.has("ASCII_Input_Filename")) throw std::runtime_error("two input files have been defined");
            } else if (PM.has("ASCII_Input_Filename")) {
                isAsc = true;
            }
        }
        const bool isForcedEk = PM.has("InputEkFilename") && PM.has("Last_K") && PM.has("Final_Forced_Time");
        const long last_input_rad = PM.integer("Last_K", 0);
        const bool isNullify = isForcedEk && PM.integer("Nullify_MissingK", 0);
        double Final_Forced_Time = PM.real("Final_Forced_Time", 0.0);
        if (isForcedEk) Final_Forced_Time = Final_Forced_Time > 0.0 ? Final_Forced_Time : Final_Time;
        int Lx = 1, Ly = 1, Lz = 1;
        bool isNotCubic = false;
        if ((PM.has("Lx") && PM.integer("Lx") > 1) || (PM.has("Ly") && PM.integer("Ly") > 1) || (PM.has("Lz") && PM.integer("Lz") > 1)) {
            Lx = (int)PM.integer("Lx", 1); Ly = (int)PM.integer("Ly", 1); Lz = (int)PM.integer("Lz", 1);
            const int l = lcm2(Lx, lcm2(Ly, Lz));
            Lx = l / Lx; Ly = l / Ly; Lz = l / Lz;
            if (Lx > 1 || Ly > 1 || Lz > 1) {
                isNotCubic = true;
                Mx = (Lx > 1 && Mx % 2 != 0) ? Mx + 1 : Mx;
                My = (Ly > 1 && My % 2 != 0) ? My + 1 : My;
                Mz = (Lz > 1 && Mz % 2 != 0) ? Mz + 1 : Mz;
                auto ex = [](long n, int L) { return (n % 2 == 0) ? n * L : (n - 1) * L + 1; };
                Nx = ex(Nx, Lx); Ny = ex(Ny, Ly); Nz = ex(Nz, Lz);
                Mx = ex(Mx, Lx); My = ex(My, Ly); Mz = ex(Mz, Lz);
                if (isAsc || isBin) { Nx_file = ex(Nx_file, Lx); Ny_file = ex(Ny_file, Ly); Nz_file = ex(Nz_file, Lz); }
            } else {
                Lx = Ly = Lz = 1;
            }
        }
        const double ox = PM.real("omega_x", 0.0), oy = PM.real("omega_y", 0.0), oz = PM.real("omega_z", 0.0);
        const bool isRotating = std::fabs(ox) > 0.0 || std::fabs(oy) > 0.0 || std::fabs(oz) > 0.0;
        const double C_Smag = PM.real("C_Smag", 0.0);
        const bool isLES = C_Smag > 0.0;
        const double C_At = PM.real("C_At", 1.0);
        const bool isSA = PM.integer("SelfAdaptiveTimestep", 0) == 1;
        const bool isCCC = PM.integer("ComplexConjugateCorrection", 0) == 1;
        const long Energy_Cascade = PM.integer("Energy_Cascade", 0), Energy_Cascade_Freq = PM.integer("Energy_Cascade_Freq", 0);
        const long VelGradInv = PM.integer("VelGradInv_Spectra", 0), VelGradInv_Freq = PM.integer("VelGradInv_Spectra_Freq", 0);
        const long Vel_Phys = PM.integer("Velocity_Physical", 0), Vel_Phys_Freq = PM.integer("Velocity_Phys_Freq", 0);
        const long Vel_Four = PM.integer("Velocity_Fourier", 0), Vel_Four_Freq = PM.integer("Velocity_Four_Freq", 0);
        const std::string spec_dir = PM.word("Spectra_Folder", ".");
        const std::string phys_dir = PM.word("Velocity_Phys_Folder", ".");
        const std::string four_dir = PM.word("Velocity_Four_Folder", ".");

        Options opt;
        const std::string prec = PM.word("Precision", "fp64");
        opt.precision = (prec == "fp32split" || prec == "fp32") ? SPNS_PREC_FP32_SPLIT : SPNS_PREC_FP64;
        opt.fixed_dt = PM.real("Fixed_Timestep", 0.0);
        opt.delta = PM.real("Delta", 0.0);
        opt.force_on = PM.has("Constant_Power");
        opt.force_eps = PM.real("Constant_Power", 0.1);
        opt.force_kmin = (int)PM.integer("Forcing_Kmin", 1);
        opt.force_kmax = (int)PM.integer("Forcing_Kmax", 2);
        opt.random_init = PM.has("Random_Seed");
        opt.seed = (unsigned long long)PM.integer("Random_Seed", 0);
        opt.k0 = PM.real("Random_k0", 4.0);
        opt.k41_ranks = (int)PM.integer("K41_Ranks", 1);
        opt.gradient_invariants = VelGradInv == 1;
        opt.keep_step_start = Vel_Phys == 1 && Vel_Phys_Freq > 0;  // in-loop dumps write the step-start field
        opt.qa_check = qa;  // the QA build checks the energy before every step (src/HIT_LES.cpp:10-12)
        const long max_iter = PM.integer("Max_Iterations", 0);
        static const char* const inv_names[5] = {"QG", "QS", "RG", "RS", "V2"};
        char inv_fname[5][1024];
        const char* inv_ptr[5] = {inv_fname[0], inv_fname[1], inv_fname[2], inv_fname[3], inv_fname[4]};

        HIT hit(Nx, Ny, Nz, Mx, My, Mz, nu, C_Smag, C_At, Lx, Ly, Lz, ox, oy, oz,
                isAsc ? PM.word("ASCII_Input_Filename").c_str() : nullptr,
                isBin ? PM.word("Binary_Input_Filename").c_str() : nullptr, Nx_file, Ny_file, Nz_file, isReLambda, isBin,
                isAsc, isNotCubic, isRotating, isLES, isCCC, isSA, opt);
        std::vector<double> Forced_Ek;
        if (isForcedEk) Forced_Ek = read_ek(PM.word("InputEkFilename"), (int)last_input_rad);
        if (isReLambda) {  // src/MAIN.cpp:337-357
            if (isForcedEk) {
                double et = 0.0, ot = 0.0;
                for (long K = 1; K <= last_input_rad; K++) { et += Forced_Ek[K]; ot += (2 * K * K * Forced_Ek[K]); }
                nu = et / ReLambda * std::sqrt(20.0 / 3.0 / ot);
            } else if (isBin || isAsc) {
                nu = hit.getEk_init_file() / ReLambda * std::sqrt(20.0 / 3.0 / hit.getEnstrophy_init_file());
            } else {
                nu = 1.0e-3;
            }
            hit.setnu(nu);
            hit.Recalculate_TimeStep();
        }
        std::printf("\n### SpNS pseudospectral HIT on MI355X: Nx %ld Ny %ld Nz %ld, Mx %ld My %ld Mz %ld, %s%s%s\n\n", Nx, Ny,
                    Nz, Mx, My, Mz, opt.precision ? "FP32-split" : "FP64", isLES ? ", Smagorinsky" : "",
                    isRotating ? ", rotating" : "");
        char fname[1024];
        long iter = 0, Ek_file_iter = 0, VelFour_file_iter = 0, VelPhys_file_iter = 0, Inv_file_iter = 0;
        const auto t0 = std::chrono::steady_clock::now();
        while (hit.gettime() < Final_Time) {  // src/MAIN.cpp:413
            hit.New_Fractional_Step_Method();
            if (iter == 0 && isReLambda && !isForcedEk && !isBin && !isAsc) {
                hit.Forcing_Reynolds_Lambda(ReLambda);
                hit.Recalculate_TimeStep();
            }
            if (isForcedEk && hit.gettime() < Final_Forced_Time)
                hit.Forcing_Energy_Cascade(Forced_Ek.data(), last_input_rad, isNullify);
            if (isReLambda_iter && iter % ReLambda_Freq == 0) {
                hit.Forcing_Reynolds_Lambda(ReLambda);
                hit.Recalculate_TimeStep();
            }
            iter++;
            if (qa || iter % 1000 == 0) {
                hit.Recalculate_Reynolds_Lambda();
                std::printf("Time: %f,    iter: %8ld,    Dt: %e,    Ek: %e,    Om: %e,    ReLambda: %8.3f\n", hit.gettime(), iter,
                            hit.getAt(), hit.getEk_Tot(), hit.getOk_Tot(), hit.getReLambda());
            }
            if (Energy_Cascade == 1 && Energy_Cascade_Freq > 0 && (iter + 1) % Energy_Cascade_Freq == 0) {
                std::snprintf(fname, sizeof fname, "%s/Energy_Cascade_%08ld.dat", spec_dir.c_str(), Ek_file_iter++);
                hit.Recalculate_Energy_Cascade(fname);
            }
            if (VelGradInv == 1 && VelGradInv_Freq > 0 && (iter + 1) % VelGradInv_Freq == 0) {  // src/MAIN.cpp:460-469
                for (int f = 0; f < 5; f++)
                    std::snprintf(inv_fname[f], sizeof inv_fname[f], "%s/%s_%08ld.dat", spec_dir.c_str(), inv_names[f], Inv_file_iter);
                Inv_file_iter++;
                hit.Recalculate_Invariants_Distribution(inv_ptr);
            }
            if (Vel_Four == 1 && Vel_Four_Freq > 0 && (iter + 1) % Vel_Four_Freq == 0) {
                std::snprintf(fname, sizeof fname, "%s/Velocity_Fourier_%08ld.bin", four_dir.c_str(), VelFour_file_iter++);
                hit.FourierVelocity_to_BinaryFile(fname);
            }
            if (Vel_Phys == 1 && Vel_Phys_Freq > 0 && iter % Vel_Phys_Freq == 0) {
                std::snprintf(fname, sizeof fname, "%s/Velocity_%08ld.bin", phys_dir.c_str(), VelPhys_file_iter++);
                hit.RealVelocity_to_BinaryFile(fname);
            }
            if (max_iter > 0 && iter >= max_iter) break;
        }
        const double loop_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (Energy_Cascade == 1) {
            std::snprintf(fname, sizeof fname, "%s/Energy_Cascade.dat", spec_dir.c_str());
            hit.Recalculate_Energy_Cascade(fname);
        }
        if (VelGradInv == 1) {  // src/MAIN.cpp:496-503
            for (int f = 0; f < 5; f++)
                std::snprintf(inv_fname[f], sizeof inv_fname[f], "%s/%s.dat", spec_dir.c_str(), inv_names[f]);
            hit.Recalculate_Invariants_Distribution(inv_ptr);
        }
        if (Vel_Four == 1) {
            std::snprintf(fname, sizeof fname, "%s/Velocity_Fourier.bin", four_dir.c_str());
            hit.FourierVelocity_to_BinaryFile(fname);
        }
        if (Vel_Phys == 1) {  // src/MAIN.cpp:510-514: antitransform first, then the current state's field
            hit.Recalculate_Velocity_Antitransform();
            std::snprintf(fname, sizeof fname, "%s/Velocity.bin", phys_dir.c_str());
            hit.RealVelocity_to_BinaryFile(fname);
        }
        hit.Recalculate_Reynolds_Lambda();
        std::printf("\nFinal time-steps: %ld\n", iter);
        std::printf("Final time:       %f\n", hit.gettime());
        std::printf("Final energy:     %f\n", hit.getEk_Tot());
        std::printf("Final ReLambda:   %.3f\n", hit.getReLambda());
        std::printf("\nLoop time:  %f\nTime/iter:  %f\n\n", loop_s, iter ? loop_s / iter : 0.0);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "spns_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
