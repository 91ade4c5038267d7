// TotalSpin/src/randNum_generator.cpp written against ed/driver_io.hpp: the same options
// (-L= -Nsam= -path=) and the same file (path/randNum_Kagome-L<L>.dat, header line, rows of L
// uniform(-1, 1) numbers, appended when the file exists).  One extension: -seed= makes the file
// reproducible (the reference always seeds from std::random_device, which -seed=-1 keeps).
#include <random>
#include <string>
#include <vector>

#include "ed/driver_io.hpp"

using namespace ed;

int main(int argc, char *argv[])
{
    std::vector<pair_t> options, defaults;
    options.push_back(pair_t("L", "Size of system"));
    options.push_back(pair_t("Nsam", "# of disorder samples"));
    options.push_back(pair_t("path", "directory to load and save files"));
    options.push_back(pair_t("seed", "mt19937 seed (-1: std::random_device)"));
    defaults.push_back(pair_t("Nsam", "1"));
    defaults.push_back(pair_t("path", "."));
    defaults.push_back(pair_t("seed", "-1"));
    argsparse parser(argc, argv, options, defaults);
    const int Nsam = parser.find<int>("Nsam"), L = parser.find<int>("L");
    const long seed = parser.find<long>("seed");
    const unsigned s = seed < 0 ? std::random_device{}() : (unsigned)seed;
    randnum_file_write(parser.find<>("path") + "/randNum_Kagome-L" + std::to_string(L) + ".dat", L, Nsam, s);
    return 0;
}
