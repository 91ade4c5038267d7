// The TotalSpin driver TotalSpinExcited_Kagome_randomHop.cu (TotalSpin/src/, main at :21-176)
// written against this library's C++ interface: the same command line (-U= -Q=up,dn -X= -Nsam=
// -dev= -nev= -path= -xpath=), the same randNum replay and the same output on stdout (the
// reference's default: its __OUTFILE__ is not defined).  For each disorder sample: the random-hop
// 24-site Kagome model (t_ij = -r_i r_j, on-site -2 r_i^2, r_i = 1 + X x_i), the nev lowest
// eigenpairs by cuSpectra::SymEigsSolver with 4 (nev + 1) basis vectors (SmallestAlge, 1000
// iterations, tol 1e-10), and per eigenvector (S^+S^- + S^-S^+)/2 and S(S+1) from
// TotalSpin::Hubbard.  The eigenvalue line prints the values separated by single spaces (Eigen's
// row-vector printing pads them to a common width).
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "ed/driver_io.hpp"
#include "ed/ed.hpp"

using namespace ed;

static const int N = 24;

// KagomeLattice24 with random hopping (TotalSpinExcited_Kagome_randomHop.cu:206-226)
static void KagomeLattice24(H_TBq &Hup, H_TBq &Hdn, const double *r)
{
    static const int first[48] = {0, 1, 6, 8, 15, 16, 21, 23, 3, 4, 9, 11, 12, 13, 18, 20,
                                  3, 2, 6, 7, 0, 5, 9, 10, 15, 14, 18, 19, 12, 17, 21, 22,
                                  1, 2, 4, 5, 7, 8, 10, 11, 13, 14, 16, 17, 19, 20, 22, 23};
    static const int second[48] = {1, 6, 8, 15, 16, 21, 23, 0, 4, 9, 11, 12, 13, 18, 20, 3,
                                   2, 6, 7, 12, 5, 9, 10, 15, 14, 18, 19, 0, 17, 21, 22, 3,
                                   2, 4, 5, 1, 8, 10, 11, 7, 14, 16, 17, 13, 20, 22, 23, 19};
    for (int b = 0; b < 48; ++b) {
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
    options.push_back(pair_t("nev", "# of eigen pairs"));
    options.push_back(pair_t("path", "directory to load and save files"));
    options.push_back(pair_t("xpath", "directory to load disorders file"));
    defaults.push_back(pair_t("X", "0"));
    defaults.push_back(pair_t("Nsam", "1"));
    defaults.push_back(pair_t("path", "."));
    defaults.push_back(pair_t("xpath", "None"));
    argsparse parser(argc, argv, options, defaults);

    const int Nsam = parser.find<int>("Nsam"), nev = parser.find<int>("nev"), dev = parser.find<int>("dev");
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

        using ModelType = cuSpectra::Hubbard<double>;
        using SolverType = cuSpectra::SymEigsSolver<ModelType>;

        H_TBq HqUP(N, Qup), HqDN(N, Qdn);
        const int64_t dimUP = HqUP.count, dimDN = HqDN.count, D = dimUP * dimDN;
        Thrust_Lin<double> LinAlg(dimUP, dimDN);
        std::cout << "# Reduced Hilbert-space dim : " << D << std::endl;
        std::cout << "#       seq{ E , (S^+S^- + S^-S^+)/2 , S(S+1) }    " << std::endl;
        std::cout.precision(10);

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
            KagomeLattice24(HqUP, HqDN, randomSites.data());
            ModelType Hmat(HqUP, HqDN, U);
            TotalSpin::Hubbard Smat(HqUP, HqDN);

            SolverType eigs(Hmat, nev, 4 * (nev + 1));
            eigs.init();
            eigs.compute(cuSpectra::SortRule::SmallestAlge, 1000, 1e-10, cuSpectra::SortRule::SmallestAlge);
            std::vector<double> eval;
            std::vector<double> evec;
            if (eigs.info() == cuSpectra::CompInfo::Successful) {
                eval = eigs.eigenvalues();
                evec = eigs.eigenvectors();
                std::cout << "#    Eigenvalues : ";
                for (int i = 0; i < nev; ++i) std::cout << (i ? " " : "") << eval[i];
                std::cout << std::endl;
            } else {
                // the reference then reads empty Eigen objects; here the sample is reported and skipped
                std::cout << "#    Eigenvalues : not converged" << std::endl;
                HqUP.clear();
                HqDN.clear();
                continue;
            }
            std::vector<double> Sxy(nev), Sop2(nev);
            for (int i = 0; i < nev; ++i) {
                device_vector<double> V(std::vector<double>(evec.begin() + (size_t)i * D,
                                                            evec.begin() + (size_t)(i + 1) * D)),
                    f((size_t)D);
                Smat.Hubbard_mv(V, f);
                Sxy[i] = LinAlg.thrust_Dotprod(V, f);
                Sop2[i] = 0.25 * (Qup - Qdn) * (Qup - Qdn) + Sxy[i];
            }
            for (int i = 0; i < nev; ++i) std::cout << "\t" << eval[i] << "\t" << Sxy[i] << "\t" << Sop2[i];
            std::cout << std::endl;
            HqUP.clear();
            HqDN.clear();
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "totalspin_excited: %s\n", e.what());
        return 1;
    }
    return 0;
}
