// The DrudeWeight driver DrudeWeight/src/drude_v2_cal_Kagome_randomHop.cu (main at :59-181,
// drude_weight_Jx at :183-220) written against this library's C++ interface, for the 12-site
// Kagome cluster (KagomeLattice12 / KagomeKinetic12 / KagomeCurrent12, :790-870; VOL = 8 sqrt 3):
// per disorder sample r_i = 1 + X x_i, the random-hop ground state by Lanczos, then the Kubo
// Drude weights D_xx and D_yy from the continued-fraction correlator of J_x, J_y (8000 omega,
// step 0.004, delta 0.01, cutoff 0.01) and the kinetic terms K_x, K_y, written to
// path/DrudeWeight_v2_Kagome-L12U<U>Q<Q>X<X>.dat as "E  Dxx  Dyy  D".  The reference compiles
// one cluster at a time (N = 24 there); -seed= (an extension) replaces std::random_device.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "ed/driver_io.hpp"
#include "ed/ed.hpp"

using namespace ed;

static const int N = 12;
static const int NOMEGA = 8000;
static const double OMEGA_STEP = 0.004, DELTA = 0.01, OMEGA_CUTOFF = 0.01;
static const double VOL = 8. * std::sqrt(3.); // 12 sites

// KagomeLattice12 bonds; the x projections of KagomeKinetic12 / KagomeCurrent12 (1 for the 8
// horizontal bonds, then d_x = 1/2), the y bonds with d_y = sqrt(3)/2
static const int B1[24] = {0, 1, 3, 4, 6, 7, 9, 10, 0, 2, 6, 8, 3, 5, 9, 11, 2, 1, 8, 7, 5, 4, 11, 10};
static const int B2[24] = {1, 3, 4, 0, 7, 9, 10, 6, 2, 6, 8, 0, 5, 9, 11, 3, 1, 11, 7, 5, 4, 8, 10, 2};
static const int Y1[16] = {0, 2, 6, 8, 3, 5, 9, 11, 1, 11, 7, 5, 4, 8, 10, 2};
static const int Y2[16] = {2, 6, 8, 0, 5, 9, 11, 3, 2, 1, 8, 7, 5, 4, 11, 10};

static void KagomeLattice12(H_TBq &H, const double *r)
{
    for (int b = 0; b < 24; ++b) H.Make_H(B1[b], B2[b], -r[B1[b]] * r[B2[b]]);
    for (int i = 0; i < N; ++i) H.Make_N(i, -2.0 * r[i] * r[i]);
}
static void KagomeKinetic12(H_TBq &Kx, H_TBq &Ky, const double *r)
{
    for (int b = 0; b < 24; ++b) Kx.Make_H(B1[b], B2[b], -r[B1[b]] * r[B2[b]] * (b < 8 ? 1.0 : 0.25));
    for (int b = 0; b < 16; ++b) Ky.Make_H(Y1[b], Y2[b], -r[Y1[b]] * r[Y2[b]] * 0.75);
}
static void KagomeCurrent12(H_TBq &Jx, H_TBq &Jy, const double *r)
{
    for (int b = 0; b < 24; ++b) Jx.Make_J(B1[b], B2[b], r[B1[b]] * r[B2[b]] * (b < 8 ? 1.0 : 0.5));
    for (int b = 0; b < 16; ++b) Jy.Make_J(Y1[b], Y2[b], r[Y1[b]] * r[Y2[b]] * (std::sqrt(3.) / 2.));
}

// drude_weight_Jx (drude_v2_cal_Kagome_randomHop.cu:183-220)
static double drude_weight_Jx(Lanczos<double> &lanc, const Hubbard &Hmat, const Hubbard &Jxmat, const Hubbard &Kxmat,
                              double eval, const device_vector<double> &evec)
{
    std::vector<std::complex<double>> omega(NOMEGA), Lambda;
    for (int i = 0; i < NOMEGA; ++i) omega[i] = std::complex<double>(OMEGA_STEP * (i + 1), DELTA);
    compute_sigma_omega(lanc, Hmat, Jxmat, eval, evec, omega, Lambda);
    double vertex = 0.0;
    for (int i = 0; i < NOMEGA; ++i) {
        const double sigma = -1.0 / (M_PI * omega[i].real()) * Lambda[i].imag();
        if (omega[i].real() > OMEGA_CUTOFF) vertex += sigma * OMEGA_STEP;
    }
    device_vector<double> KV(evec.size());
    Kxmat.Hubbard_mv(evec, KV);
    const double kinetic = lanc.thrust_Dotprod(evec, KV);
    return (-kinetic - 2.0 * vertex) / VOL;
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
    options.push_back(pair_t("seed", "mt19937 seed (-1: std::random_device)"));
    defaults.push_back(pair_t("X", "0"));
    defaults.push_back(pair_t("Nsam", "1"));
    defaults.push_back(pair_t("path", "."));
    defaults.push_back(pair_t("seed", "-1"));
    argsparse parser(argc, argv, options, defaults);
    const int Nsam = parser.find<int>("Nsam"), dev = parser.find<int>("dev");
    const double U = parser.find<double>("U"), X = parser.find<double>("X");
    const auto Q = parser.mfind<int>("Q");
    const int Qup = Q[0], Qdn = Q[1];
    const long seed = parser.find<long>("seed");
    parser.print(std::cout);
    try {
        hip_check(hipSetDevice(dev), "hipSetDevice");
        std::random_device rn;
        std::mt19937 generator(seed < 0 ? rn() : (unsigned)seed);
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        std::vector<double> randomSites(N);
        H_TBq H0UP(N, Qup), H0DN(N, Qdn), KxUP(N, Qup), KxDN(N, Qdn), KyUP(N, Qup), KyDN(N, Qdn), JxUP(N, Qup),
            JxDN(N, Qdn), JyUP(N, Qup), JyDN(N, Qdn);
        const int64_t D = H0UP.count * H0DN.count;
        std::cout << "# Reduced Hilbert space : " << D << std::endl;
        device_vector<double> f((size_t)D), V((size_t)D, 0.0);
        double E = 0.0;
        const std::string filename = parser.find<>("path") + "/DrudeWeight_v2_Kagome-L" + std::to_string(N) + "U" +
                                     parser.find<>("U") + "Q" + parser.find<>("Q") + "X" + parser.find<>("X") + ".dat";
        DatFile out(filename, "#       E       Dxx        Dyy        D      ");
        for (int sam = 0; sam < Nsam; ++sam) {
            for (int i = 0; i < N; ++i) randomSites[i] = 1.0 + X * dist(generator);
            KagomeLattice12(H0UP, randomSites.data());
            KagomeLattice12(H0DN, randomSites.data());
            KagomeKinetic12(KxUP, KyUP, randomSites.data());
            KagomeKinetic12(KxDN, KyDN, randomSites.data());
            KagomeCurrent12(JxUP, JyUP, randomSites.data());
            KagomeCurrent12(JxDN, JyDN, randomSites.data());
            Hubbard Hmat(H0UP, H0DN, U), Kxmat(KxUP, KxDN), Jxmat(JxUP, JxDN), Kymat(KyUP, KyDN), Jymat(JyUP, JyDN);
            Lanczos<double> lanc(H0UP.count, H0DN.count);
            lanc.random_vector(f);
            lanc.eigen(Hmat, f, E, V);
            const double dxx = drude_weight_Jx(lanc, Hmat, Jxmat, Kxmat, E, V);
            const double dyy = drude_weight_Jx(lanc, Hmat, Jymat, Kymat, E, V);
            out.row({E, dxx, dyy, (dxx + dyy) / 2.});
            std::cout << "\r# ---- " << std::setw(4) << sam + 1 << "/" << std::setw(4) << Nsam << std::flush;
            for (H_TBq *h : {&H0UP, &H0DN, &KxUP, &KxDN, &KyUP, &KyDN, &JxUP, &JxDN, &JyUP, &JyDN}) h->clear();
        }
        std::cout << std::endl;
    } catch (const Error &e) {
        std::fprintf(stderr, "drude_weight: %s\n", e.what());
        return 1;
    }
    return 0;
}
