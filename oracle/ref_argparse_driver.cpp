// ref_argparse_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Runs the reference's own command-line parser, MBL_test/include/argparse.h (struct argsparse),
// included from the reference tree at build time (oracle/Makefile, target `ref`; g++, host
// only), on the option set of the TotalSpin drivers (ConservedTotalSpin_Kagome_randomHop.cu:21-
// 35), and prints what a driver reads back.  tests/test_driver_io.py runs it and this library's
// ed::argsparse (tests/cpp_host/driver_io_check.cpp) on the same argument lists and compares
// stdout, stderr and exit status.
#include <array>
#include "argparse.h"

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
