// The TotalSpin driver ConservedTotalSpin_Kagome_randomHop.cu (TotalSpin/src/, main at :19-138)
// written against this library's C++ interface: the same command line (-U= -Q=up,dn -X= -Nsam=
// -dev= -path= -xpath=), the same randNum replay and the same .dat output, with the file output
// switched on as the other drivers do (#define __OUTFILE__ there).  For each disorder sample:
// the random-hop 12-site Kagome model (t_ij = -r_i r_j, on-site -2 r_i^2, r_i = 1 + X x_i),
// the Lanczos ground state, S(S+1) from TotalSpin::Hubbard.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "ed/driver_io.hpp"
#include "ed/ed.hpp"

#define ED_OUTFILE

using namespace ed;

static const int N = 12;

// KagomeLattice12 with random hopping (ConservedTotalSpin_Kagome_randomHop.cu:140-156)
static void KagomeLattice12(H_TBq &Hup, H_TBq &Hdn, const double *r)
{
    static const int first[24] = {0, 1, 3, 4, 6, 7, 9, 10, 0, 2, 6, 8, 3, 5, 9, 11, 2, 1, 8, 7, 5, 4, 11, 10};
    static const int second[24] = {1, 3, 4, 0, 7, 9, 10, 6, 2, 6, 8, 0, 5, 9, 11, 3, 1, 11, 7, 5, 4, 8, 10, 2};
    for (int b = 0; b < 24; ++b) {
        Hup.Make_H(first[b], second[b], -r[first[b]] * r[second[b]]);
        Hdn.Make_H(first[b], second[b], -r[first[b]] * r[second[b]]);
    }
    for (int i = 0; i < N; ++i) {
        Hup.Make_N(i, -2.0 * r[i] * r[i]);
        Hdn.Make_N(i, -2.0 * r[i] * r[i]);
    }
}

int main(int argc, char *argv[])
{
    std::vector<pair_t> options, defaults;
    options.push_back(pair_t("U", "interaction"));
    options.push_back(pair_t("Q", "# of electrons"));
    options.push_back(pair_t("X", "disorder strength (0<= X <1)"));
    options.push_back(pair_t("Nsam", "# of disorder samples"));
    options.push_back(pair_t("dev", "device number"));
    options.push_back(pair_t("path", "directory to load and save files"));
    options.push_back(pair_t("xpath", "directory to load disorders file"));
    defaults.push_back(pair_t("X", "0"));
    defaults.push_back(pair_t("Nsam", "1"));
    defaults.push_back(pair_t("path", "."));
    defaults.push_back(pair_t("xpath", "None"));
    argsparse parser(argc, argv, options, defaults);

    const int Nsam = parser.find<int>("Nsam"), dev = parser.find<int>("dev");
    const double U = parser.find<double>("U"), X = parser.find<double>("X");
    const auto Q = parser.mfind<int>("Q");
    const int Qup = Q[0], Qdn = Q[1];
    parser.print(std::cout);
    try {
        hip_check(hipSetDevice(dev), "hipSetDevice");
        std::random_device rn;
        std::mt19937 generator(rn());
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        std::vector<double> randomSites(N);

        H_TBq HqUP(N, Qup), HqDN(N, Qdn);
        const int64_t D = HqUP.count * HqDN.count;
        std::cout << "# Reduced Hilbert-space dim : " << D << std::endl;
        device_vector<double> f((size_t)D), V((size_t)D, 0.0);
        double E = 0.0;

        const std::string filename = parser.find<>("path") + "/ConservedTotalSpin_Kagome-L" + std::to_string(N) + "U" +
                                     parser.find<>("U") + "Q" + parser.find<>("Q") + "X" + parser.find<>("X") + ".dat";
#ifdef ED_OUTFILE
        DatFile out(filename, "#       E       (S^+S^- + S^-S^+)/2       S(S+1)    ");
#else
        std::cout << "#       E        (S^+S^- + S^-S^+)/2        S(S+1)     " << std::endl;
        std::cout.precision(10);
#endif
        RandNumReplay replay;
        const bool from_file = parser.find<>("xpath") != "None";
        if (from_file &&
            !replay.open(parser.find<>("xpath") + "/randNum_Kagome-L" + std::to_string(N) + ".dat"))
            std::cout << "There is not the file!" << std::endl;

        for (int sam = 0; sam < Nsam; ++sam) {
            if (from_file) {
                std::vector<double> x;
                replay.next(N, x);
                for (int i = 0; i < N; ++i) randomSites[i] = 1.0 + X * x[i];
            } else {
                for (int i = 0; i < N; ++i) randomSites[i] = 1.0 + X * dist(generator);
            }
            KagomeLattice12(HqUP, HqDN, randomSites.data());
            Hubbard Hmat(HqUP, HqDN, U);
            TotalSpin::Hubbard Smat(HqUP, HqDN);
            Lanczos<double> lanc(HqUP.count, HqDN.count);
            lanc.random_vector(f);
            lanc.eigen(Hmat, f, E, V);
            Smat.Hubbard_mv(V, f);
            const double Sxy = lanc.thrust_Dotprod(V, f);
            const double Sop2 = 0.25 * (Qup - Qdn) * (Qup - Qdn) + Sxy;
#ifdef ED_OUTFILE
            out.row({E, Sxy, Sop2});
            std::cout << "\r# ---- " << std::setw(4) << sam + 1 << "/" << std::setw(4) << Nsam << std::flush;
#else
            std::cout << "\t" << E << "\t" << Sxy << "\t" << Sop2 << std::endl;
#endif
            HqUP.clear();
            HqDN.clear();
        }
#ifdef ED_OUTFILE
        std::cout << std::endl;
#endif
    } catch (const Error &e) {
        std::fprintf(stderr, "conserved_total_spin: %s\n", e.what());
        return 1;
    }
    return 0;
}
