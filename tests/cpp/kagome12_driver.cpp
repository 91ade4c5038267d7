// A driver written against the reference's C++ interfaces (through include/ed/ed.hpp): the
// 12-site Kagome Hubbard model at N_up = N_dn = 6, U = 4, hopping -1, following the steps of
// the DrudeWeight and TotalSpin drivers (drude_v2_cal_Kagome_randomHop.cu:129-220 and
// ConservedTotalSpin_Kagome_randomHop.cu:110-125): ground state by Lanczos, S(S+1), the Kubo
// Drude weight D_x, the 6|6 entanglement entropy, the three lowest eigenvalues from the
// cuSpectra solver and one host-array perform_op; then the same ground state on the slab layout,
// and the dense path of the MBL drivers (Hubbard::Write, cusolverDsyevd, tED, H_TBq::Write with
// Insert_1Fop / Insert_2Fop) and the LOW_MEM basis on small systems.  Prints one JSON object.
#include <complex>
#include <cstdio>
#include <vector>

#include "ed/ed.hpp"

using namespace ed;

// KagomeLattice12 bonds (ConservedTotalSpin_Kagome_randomHop.cu:140-156); the first 8 are the
// horizontal A-B bonds (x projection 1), the other 16 have x projection 1/2
static const int kFirst[24] = {0, 1, 3, 4, 6, 7, 9, 10, 0, 2, 6, 8, 3, 5, 9, 11, 2, 1, 8, 7, 5, 4, 11, 10};
static const int kSecond[24] = {1, 3, 4, 0, 7, 9, 10, 6, 2, 6, 8, 0, 5, 9, 11, 3, 1, 11, 7, 5, 4, 8, 10, 2};

static void KagomeLattice12(H_TBq &Hq)
{
    for (int b = 0; b < 24; ++b) Hq.Make_H(kFirst[b], kSecond[b], -1.0);
}
static void KineticX(H_TBq &Hq) // K_x: d_x^2 weights (drude_v2_cal_Kagome_randomHop.cu:810-839)
{
    for (int b = 0; b < 24; ++b) Hq.Make_H(kFirst[b], kSecond[b], -(b < 8 ? 1.0 : 0.25));
}
static void CurrentX(H_TBq &Hq) // J_x: d_x weights (drude_v2_cal_Kagome_randomHop.cu:841-870)
{
    for (int b = 0; b < 24; ++b) Hq.Make_J(kFirst[b], kSecond[b], b < 8 ? 1.0 : 0.5);
}

int main()
{
    const int N = 12, QUP = 6, QDN = 6;
    const double U = 4.0, VOL = 8.0 * std::sqrt(3.0);
    try {
        H_TBq HqUP(N, QUP), HqDN(N, QDN);
        KagomeLattice12(HqUP);
        KagomeLattice12(HqDN);
        Hubbard Hub(HqUP, HqDN, U);
        const int64_t D = Hub.dimUP * Hub.dimDN;

        Lanczos<double> lanc(Hub.dimUP, Hub.dimDN);
        device_vector<double> f(D), evec(D);
        double eval = 0.0;
        lanc.random_vector(f);
        lanc.eigen(Hub, f, eval, evec);

        TotalSpin::Hubbard S2(HqUP, HqDN);
        device_vector<double> s2v(D);
        S2.Hubbard_mv(evec, s2v);
        const double Sxy = lanc.thrust_Dotprod(evec, s2v);
        const double SS = 0.25 * (QUP - QDN) * (QUP - QDN) + Sxy;

        // Kubo Drude weight (drude_weight_Jx)
        H_TBq JUP(N, QUP), JDN(N, QDN), KUP(N, QUP), KDN(N, QDN);
        CurrentX(JUP);
        CurrentX(JDN);
        KineticX(KUP);
        KineticX(KDN);
        Hubbard Jx(JUP, JDN, 0.0), Kx(KUP, KDN, 0.0);
        const int NOMEGA = 8000;
        std::vector<std::complex<double>> omega(NOMEGA), Lambda;
        for (int k = 0; k < NOMEGA; ++k) omega[k] = std::complex<double>(0.004 * (k + 1), 0.01);
        Lanczos<double> lanctri(Hub.dimUP, Hub.dimDN);
        compute_sigma_omega(lanctri, Hub, Jx, eval, evec, omega, Lambda);
        double vertex = 0.0;
        for (int k = 0; k < NOMEGA; ++k) {
            const double sigma = -1.0 / (M_PI * omega[k].real()) * Lambda[k].imag();
            if (omega[k].real() > 0.01) vertex += sigma * 0.004;
        }
        device_vector<double> kv(D);
        Kx.Hubbard_mv(evec, kv);
        const double kinetic = lanc.thrust_Dotprod(evec, kv);
        const double Dw = (-kinetic - 2.0 * vertex) / VOL;

        // 6|6 entanglement entropy (Asite = {0..5})
        int Asite[6] = {0, 1, 2, 3, 4, 5};
        EEmatHubAB_GPU<double> EE(Asite, 6, HqUP, HqDN);
        const auto ent = EE.entropy(evec);

        // three lowest eigenvalues (cuSpectra solver) and one host-array perform_op
        cuSpectra::Hubbard<double> op(HqUP, HqDN, U);
        cuSpectra::SymEigsSolver<cuSpectra::Hubbard<double>> eigs(op, 3, 16);
        eigs.init();
        const int nconv = eigs.compute(cuSpectra::SortRule::SmallestAlge, 1000, 1e-12);
        std::vector<double> xh = evec.to_host(), yh(D);
        op.perform_op(xh.data(), yh.data());
        double rq = 0.0;
        for (int64_t i = 0; i < D; ++i) rq += xh[i] * yh[i];

        // quench as in the MBL drivers (Hubbard_EE_test.cu:44-50, 76-83) with Krylov<C>
        // (HubMatrix_GPU.cuh:619-728): the product state up = sites 0-5, dn = sites 6-11, evolved
        // by dt = 0.2 in an 80-step Krylov space, then its 6|6 entanglement entropy
        std::vector<std::complex<double>> ph((size_t)D, 0.0);
        ph[(size_t)HqUP.address[0x03f] * Hub.dimDN + HqDN.address[0xfc0]] = 1.0;
        device_vector<std::complex<double>> psi0(ph), psi((size_t)D);
        Krylov<double> kr(Hub.dimUP, Hub.dimDN);
        kr.maxN = 80;
        kr.Make_tridiagonal_eigen(Hub, psi0);
        kr.Krylov_time_evol(Hub, psi0, psi, 0.2);
        EEmatHubAB_GPU<std::complex<double>> EEq(Asite, 6, HqUP, HqDN);
        const auto entq = EEq.entropy(psi);
        const double qnorm = kr.norm(psi);

        // the same ground state with the recurrence on the slab layout
        Hub.slab_layout();
        Lanczos<double> lanc_s(Hub.dimUP, Hub.dimDN);
        device_vector<double> fs(D), evs(D);
        double eval_s = 0.0;
        lanc_s.random_vector(fs);
        lanc_s.eigen(Hub, fs, eval_s, evs);
        const double ovl = lanc.thrust_Dotprod(evec, evs);
        Hub.slab_layout(false);

        // dense path (Time_evol_test.cu:73-94): 8-site ring, 4/4, U = 4
        H_TBq RU(8, 4), RD(8, 4);
        for (int i = 0; i < 8; ++i) {
            RU.Make_H(i, (i + 1) % 8, -1.0);
            RD.Make_H(i, (i + 1) % 8, -1.0);
        }
        Hubbard Ring(RU, RD, 4.0);
        const int64_t DR = Ring.dimUP * Ring.dimDN;
        device_vector<double> Hmat((size_t)(DR * DR), 0.0);
        Ring.Write(Hmat);
        tED<double> ted(Ring.dimUP, Ring.dimDN);
        ted.Make_eigen(Hmat);
        const std::vector<double> W = ted.W.to_host();
        std::vector<std::complex<double>> p0((size_t)DR, 0.0);
        const int64_t k0 = (int64_t)RU.address[0x0f] * Ring.dimDN + RD.address[0xf0];
        p0[(size_t)k0] = 1.0;
        device_vector<std::complex<double>> phi0(p0), phi((size_t)DR);
        ted.tED_time_evol(phi0, phi, 1.5);
        const std::vector<std::complex<double>> ph1 = phi.to_host();
        double tnorm = 0.0;
        for (const auto &z : ph1) tnorm += std::norm(z);
        Lanczos<double> lr(Ring.dimUP, Ring.dimDN);
        device_vector<double> fr((size_t)DR), er((size_t)DR);
        double e_ring = 0.0;
        lr.random_vector(fr);
        lr.eigen(Ring, fr, e_ring, er);

        // H_TBq::Write on the full Fock space of 5 sites with Insert_1Fop / Insert_2Fop
        H_TBq F5(5);
        F5.Insert_1Fop('+', 2, 0.5);
        F5.Insert_2Fop('-', 4, '+', 1, -0.75);
        F5.Make_N(3, 0.25);
        device_vector<double> M5(32 * 32, 0.0);
        F5.Write(M5);
        const std::vector<double> m5 = M5.to_host();
        double m5w = 0.0; // weighted checksum sum_jk (j*32 + k + 1) M[j, k]
        for (int j = 0; j < 32 * 32; ++j) m5w += (j + 1) * m5[j];

        const LOW_MEM::H_TBq_basis lm(4, 2);

        const auto &ev = eigs.eigenvalues();
        std::printf("{\"E0\": %.17g, \"nsteps\": %d, \"SS\": %.17g, \"D\": %.17g, \"kinetic\": %.17g, "
                    "\"S_vN\": %.17g, \"S_ref\": %.17g, \"n_schmidt\": %lld, \"nconv\": %d, "
                    "\"evals\": [%.17g, %.17g, %.17g], \"rayleigh_host\": %.17g, \"krylov_steps\": %d, "
                    "\"quench_norm2\": %.17g, \"quench_S_vN\": %.17g, \"E0_slab\": %.17g, \"nsteps_slab\": %d, "
                    "\"overlap_slab\": %.17g, \"ring_W\": [%.17g, %.17g, %.17g], \"ring_lanczos_E0\": %.17g, "
                    "\"ted_norm2\": %.17g, \"ted_amp0\": [%.17g, %.17g], \"fock5_checksum\": %.17g, "
                    "\"lowmem_4_2\": [%llu, %llu, %llu, %llu, %llu, %llu]}\n",
                    eval, lanc.nsteps, SS, Dw, kinetic, ent.S_vN, ent.S_ref, (long long)ent.n_schmidt, nconv, ev[0],
                    ev[1], ev[2], rq, (int)kr.a.size(), qnorm, entq.S_vN, eval_s, lanc_s.nsteps, ovl, W[0], W[1], W[2],
                    e_ring, tnorm, ph1[(size_t)k0].real(), ph1[(size_t)k0].imag(), m5w,
                    (unsigned long long)lm.state[0], (unsigned long long)lm.state[1], (unsigned long long)lm.state[2],
                    (unsigned long long)lm.state[3], (unsigned long long)lm.state[4], (unsigned long long)lm.state[5]);
    } catch (const Error &e) {
        std::fprintf(stderr, "kagome12_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
