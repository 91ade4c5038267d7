// The spinless-fermion quench driver MBL_test/src/spinless_fermion_test.cu (main at :13-84,
// Line12 at :86-118) written against this library's C++ interface: a chain of N sites with
// hopping -1, nearest-neighbour interaction U n_i n_{i+1} (Insert_2Nop) and on-site disorder
// x_coe * U(-1, 1); the dense single-species matrix (H_TBq::Write), tED with dimDN = 1, the
// product state with sites 0-5 filled, and per time step the entropy -sum s ln s of the
// A = {0..5} Schmidt values (the reference's EEmatAB + cusolverZsvd; here the two-species
// entropy with an empty down species).  Output rows "xid  t  SE" in sf_EE_U<U>_x<x>.out.
// Positional arguments as the reference (N Q U x_coe), then optionally seed Nt x_conf dev outdir.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "ed/ed.hpp"

using namespace ed;

// Line12 (spinless_fermion_test.cu:86-118) for a chain of N sites
static void LineN(H_TBq &Hq, int N, double U, double x_coe, std::mt19937 &rand,
                  std::uniform_real_distribution<double> &dist)
{
    const int t = -1;
    for (int i = 0; i + 1 < N; ++i) Hq.Make_H(i, i + 1, t);
    for (int i = 0; i + 1 < N; ++i) Hq.Insert_2Nop(i, i + 1, U);
    for (int i = 0; i < N; ++i) Hq.Make_N(i, x_coe * dist(rand));
}

int main(int argc, char *argv[])
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s N Q U x_coe [seed Nt x_conf dev outdir]\n", argv[0]);
        return 1;
    }
    const int N = std::atoi(argv[1]), Q = std::atoi(argv[2]);
    const double U = std::atof(argv[3]), x_coe = std::atof(argv[4]);
    const long seed = argc > 5 ? std::atol(argv[5]) : -1;
    const int Nt = argc > 6 ? std::atoi(argv[6]) : 500, x_conf = argc > 7 ? std::atoi(argv[7]) : 500;
    const int dev = argc > 8 ? std::atoi(argv[8]) : 1;
    const std::string outdir = argc > 9 ? argv[9] : ".";
    try {
        hip_check(hipSetDevice(dev), "hipSetDevice");
        H_TBq Hq(N, Q), Hempty(N, 0);
        const int64_t dim = Hq.count;
        const int Asite[6] = {0, 1, 2, 3, 4, 5};
        EEmatHubAB_GPU<std::complex<double>> EEfermi(Asite, 6, Hq, Hempty);
        tED<double> time_ED(dim, 1);
        device_vector<double> Hmat((size_t)(dim * dim), 0.0);
        std::vector<std::complex<double>> p0((size_t)dim, 0.0);
        p0[(size_t)Hq.address[63]] = 1.0; // state0 = 000000111111
        device_vector<std::complex<double>> psi0(p0), psi((size_t)dim);
        std::random_device rn;
        std::mt19937 rand(seed < 0 ? rn() : (unsigned)seed);
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        std::ofstream outF(outdir + "/sf_EE_U" + std::to_string((int)U) + "_x" + std::to_string((int)x_coe) + ".out");
        outF.precision(10);
        const double dt = 0.02;
        for (int i = 0; i < x_conf * Nt; ++i) {
            const int tid = i % Nt, xid = i / Nt;
            if (tid == 0) {
                LineN(Hq, N, U, x_coe, rand, dist);
                Hq.Write(Hmat);
                time_ED.Make_eigen(Hmat);
            }
            psi.resize((size_t)dim); // phi accumulates in tED_time_evol: start from zero
            time_ED.tED_time_evol(psi0, psi, std::exp((tid + 1) * dt) - 1.);
            const double SE = EEfermi.entropy(psi).S_ref;
            if (tid == 0 && xid != 0) outF << "\n" << std::endl;
            outF << xid << "\t" << std::exp((tid + 1) * dt) - 1. << "\t" << SE << std::endl;
            if (tid == Nt - 1) {
                Hq.clear();
                Hmat.resize((size_t)(dim * dim), 0.0);
                time_ED.clear();
            }
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "spinless_quench: %s\n", e.what());
        return 1;
    }
    return 0;
}
