// driver_io_check.cpp -- host program of tests/test_driver_io.py (compiled with g++, no GPU).
//
//   driver_io_check parse -U=.. -Q=..    ed::argsparse on the option set of the TotalSpin drivers
//                                        (ConservedTotalSpin_Kagome_randomHop.cu:21-35); prints
//                                        what a driver reads back, in the same format as
//                                        oracle/ref_argparse_driver.cpp prints the reference's
//   driver_io_check randnum FILE L NSAM SEED     ed::randnum_file_write
//   driver_io_check replay FILE N NSAM           ed::RandNumReplay, one line of N values per sample
//   driver_io_check dat FILE                     ed::DatFile with two rows
#include <array>
#include <cstdlib>
#include <cstring>

#include "ed/driver_io.hpp"

using namespace ed;

int main(int argc, char *argv[])
{
    if (argc >= 2 && std::strcmp(argv[1], "parse") == 0) {
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
        argv[1] = argv[0]; // the parser sees argv[1..] as the driver's own arguments
        argsparse parser(argc - 1, argv + 1, options, defaults);
        parser.print(std::cout);
        const auto Q = parser.mfind<int>("Q");
        std::cout.precision(17);
        std::cout << "U=" << parser.find<double>("U") << " X=" << parser.find<double>("X")
                  << " Nsam=" << parser.find<int>("Nsam") << " dev=" << parser.find<int>("dev")
                  << " path=" << parser.find<>("path") << " xpath=" << parser.find<>("xpath")
                  << " Qfirst=" << parser.find<int>("Q") << " Q=";
        for (int q : Q) std::cout << q << ";";
        std::cout << std::endl;
        const auto Q2 = parser.mfind<int, 2>("Q");
        std::cout << "Q2=" << Q2[0] << "," << Q2[1] << std::endl;
        return 0;
    }
    if (argc == 6 && std::strcmp(argv[1], "randnum") == 0) {
        randnum_file_write(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), (unsigned)std::atol(argv[5]));
        return 0;
    }
    if (argc == 5 && std::strcmp(argv[1], "replay") == 0) {
        RandNumReplay r;
        if (!r.open(argv[2])) std::cout << "There is not the file!" << std::endl;
        std::vector<double> x;
        std::cout.precision(17);
        for (int s = 0; s < std::atoi(argv[4]); ++s) {
            r.next(std::atoi(argv[3]), x);
            for (double v : x) std::cout << v << " ";
            std::cout << std::endl;
        }
        return 0;
    }
    if (argc == 3 && std::strcmp(argv[1], "dat") == 0) {
        DatFile d(argv[2], "#       E       (S^+S^- + S^-S^+)/2       S(S+1)    ");
        d.row({-8.123456789012345, 1.5e-12, 2.0});
        d.row({-7.0, 0.75, 0.75});
        return 0;
    }
    std::fprintf(stderr, "usage: driver_io_check parse|randnum|replay|dat ...\n");
    return 2;
}
