// The MBL quench driver MBL_test/src/Hubbard_EE_test.cu (main at :14-94) written against this
// library's C++ interface: the same positional arguments (N QUP QDN U x_coe), the same disorder
// model (hopping -1 along a chain, on-site x_coe * U(-1, 1) drawn separately for up and dn,
// Line16 at :111-134 for N = 16), the same initial product state (up = 192, dn = 768), the same
// dense path per disorder configuration (Hubbard::Write, tED::Make_eigen) and per time step
// (tED_time_evol at t = exp((tid + 1) dt) - 1, then -sum s ln s over the Schmidt values of the
// A = {0..7} bipartition), and the same output rows "xid  t  SE" with a blank line between
// configurations.  Optional extra arguments (not in the reference, which seeds from
// std::random_device, runs 500 x 100 steps on device 7 and writes ./Hubbard_model/):
//   hubbard_ee_quench N QUP QDN U x_coe [seed Nt x_conf dev outdir]
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

// Line16 (Hubbard_EE_test.cu:111-134) for a chain of N sites
static void LineN(H_TBq &Hq, int N, double x_coe, std::mt19937 &rand, std::uniform_real_distribution<double> &dist)
{
    const int t = -1;
    for (int i = 0; i + 1 < N; ++i) Hq.Make_H(i, i + 1, t);
    for (int i = 0; i < N; ++i) Hq.Make_N(i, x_coe * dist(rand));
}

int main(int argc, char *argv[])
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s N QUP QDN U x_coe [seed Nt x_conf dev outdir]\n", argv[0]);
        return 1;
    }
    const int N = std::atoi(argv[1]), QUP = std::atoi(argv[2]), QDN = std::atoi(argv[3]);
    const double U = std::atof(argv[4]), x_coe = std::atof(argv[5]);
    const long seed = argc > 6 ? std::atol(argv[6]) : -1;
    const int Nt = argc > 7 ? std::atoi(argv[7]) : 500, x_conf = argc > 8 ? std::atoi(argv[8]) : 100;
    const int dev = argc > 9 ? std::atoi(argv[9]) : 7;
    const std::string outdir = argc > 10 ? argv[10] : "./Hubbard_model";
    try {
        hip_check(hipSetDevice(dev), "hipSetDevice");
        H_TBq HqUP(N, QUP), HqDN(N, QDN);
        const int64_t dimUP = HqUP.count, dimDN = HqDN.count, dimHub = dimUP * dimDN;
        const int Asize = 8;
        const int Asite[8] = {0, 1, 2, 3, 4, 5, 6, 7};
        EEmatHubAB_GPU<std::complex<double>> EEhub(Asite, Asize, HqUP, HqDN);

        tED<double> time_ED(dimUP, dimDN);
        std::cout << "total Hmat size : " << dimHub * dimHub << std::endl;
        device_vector<double> Hmat((size_t)(dimHub * dimHub), 0.0);

        const int UPstate = 192, DNstate = 768; // up = ..0011000000, dn = 0000001100..
        std::vector<std::complex<double>> p0((size_t)dimHub, 0.0);
        p0[(size_t)((int64_t)HqUP.address[UPstate] * dimDN + HqDN.address[DNstate])] = 1.0;
        device_vector<std::complex<double>> psi0(p0), psi((size_t)dimHub);

        std::random_device rn;
        std::mt19937 rand(seed < 0 ? rn() : (unsigned)seed);
        std::uniform_real_distribution<double> dist(-1.0, 1.0);

        const std::string file_name = outdir + "/Hub_EE_L" + std::to_string(N) + "_Q" + std::to_string(QUP + QDN) +
                                      "_U" + std::to_string((int)U) + "_x" + std::to_string((int)x_coe) + ".out";
        std::ofstream outF(file_name);
        outF.precision(10);

        const double dt = 0.02;
        for (int i = 0; i < x_conf * Nt; ++i) {
            const int tid = i % Nt, xid = i / Nt;
            if (tid == 0) {
                LineN(HqUP, N, x_coe, rand, dist);
                LineN(HqDN, N, x_coe, rand, dist);
                Hubbard Hub(HqUP, HqDN, U);
                Hub.Write(Hmat);
                time_ED.Make_eigen(Hmat);
            }
            psi.resize((size_t)dimHub); // phi accumulates in tED_time_evol: start from zero
            time_ED.tED_time_evol(psi0, psi, std::exp((tid + 1) * dt) - 1.);
            const double SE = EEhub.entropy(psi).S_ref; // -sum_{s != 0} s ln s
            if (tid == 0 && xid != 0) outF << "\n" << std::endl;
            outF << xid << "\t" << std::exp((tid + 1) * dt) - 1. << "\t" << SE << std::endl;
            if (tid == Nt - 1) {
                HqUP.clear();
                HqDN.clear();
                Hmat.resize((size_t)(dimHub * dimHub), 0.0);
                time_ED.clear();
            }
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "hubbard_ee_quench: %s\n", e.what());
        return 1;
    }
    return 0;
}
