// ed/driver_io.hpp -- the driver I/O layer of the reference's programs, host C++ only:
//
//   argsparse            the command-line parser every driver builds (MBL_test/include/argparse.h):
//                        arguments "-name=value", defaults, "--help", find<T>, mfind<T>, print
//   randnum_file_write   TotalSpin/src/randNum_generator.cpp: rows of L uniform(-1, 1) numbers
//   RandNumReplay        the drivers' replay of such a file, one line per disorder sample
//                        (ConservedTotalSpin_Kagome_randomHop.cu:85-107)
//   DatFile              the append-mode, tab-separated .dat outputs with a '#' header line
//
// Argument-error behaviour follows the reference: a message on stderr and exit(1).  No
// dependence on the GPU library; a driver includes it beside ed/ed.hpp.
#pragma once

#include <array>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace ed {

using pair_t = std::pair<std::string, std::string>;

class argsparse {
  public:
    // options: (name, explanation); defaults: (name, value).  argv[k] (k >= 1) is "-name=value";
    // the first option (in list order) whose name equals the text after the leading character
    // claims the argument.  Missing options without a default, repeated options, a missing '='
    // or an empty value print an error and exit(1); "--help" prints the option list and exit(1).
    argsparse(const int argc, const char *const argv[], const std::vector<pair_t> &options,
              const std::vector<pair_t> &defaults = std::vector<pair_t>())
    {
        names_.reserve(options.size());
        for (const pair_t &o : options) names_.push_back(o.first);
        values_.assign(names_.size(), std::string());
        std::vector<bool> seen(names_.size(), false);
        for (int a = 1; a < argc; ++a)
            if (std::string(argv[a]) == "--help") help(argv[0], options, defaults);
        bool bad = false;
        for (int a = 1; a < argc; ++a) {
            const std::string arg(argv[a]);
            if (arg.empty()) continue;
            for (size_t k = 0; k < names_.size(); ++k) {
                const std::string &nm = names_[k];
                if (arg.compare(1, nm.size(), nm) != 0) continue;
                if (arg.size() <= nm.size() + 2) {
                    std::cerr << "# error(in-1) ---> Put the option correctly! : " << nm << std::endl;
                    bad = true;
                } else if (seen[k]) {
                    std::cerr << "# error(in-2) ---> The doubly occupied option is found! : " << nm << ": " << arg
                              << std::endl;
                    bad = true;
                } else if (arg[nm.size() + 1] != '=') {
                    std::cerr << "# error(in-3) ---> The symbol '=' must be in between the option and the "
                                 "argument. : "
                              << arg << std::endl;
                    bad = true;
                } else {
                    values_[k] = arg.substr(nm.size() + 2);
                    seen[k] = true;
                }
                break;
            }
        }
        for (size_t k = 0; k < names_.size(); ++k) {
            if (seen[k]) continue;
            for (const pair_t &d : defaults)
                if (d.first == names_[k]) {
                    values_[k] = d.second;
                    seen[k] = true;
                    break;
                }
            if (!seen[k]) {
                std::cerr << "# error(in) ---> The following option is missing. : " << names_[k] << std::endl;
                bad = true;
            }
        }
        if (bad) {
            std::cerr << std::endl
                      << " (hint) Type the command to the cmd line as follows: " << argv[0] << " --help" << std::endl;
            std::exit(1);
        }
    }

    // the value of one option, read with operator>> (so find<int>("Q") of "6,6" is 6)
    template <typename T = std::string>
    T find(const std::string &name) const
    {
        const std::string &text = value(name);
        T v{};
        if (!cast(text, v)) {
            std::cerr << "# error has occured in the lexical cast: " << text << std::endl;
            std::exit(1);
        }
        return v;
    }

    // a comma-separated list ("-Q=6,6"); an empty item (e.g. a trailing ',') is an error
    template <typename T = std::string>
    std::vector<T> mfind(const std::string &name) const
    {
        const std::string &all = value(name);
        std::vector<T> out;
        size_t start = 0;
        while (true) {
            const size_t comma = all.find(',', start);
            const std::string item = all.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
            if (item.empty()) {
                std::cerr << "# error has occured: remove ',' at the last part" << std::endl;
                std::exit(1);
            }
            T v{};
            if (!cast(item, v)) {
                std::cerr << "# error has occured in the lexical cast: " << all << " (option: " << name << ")"
                          << std::endl;
                std::exit(1);
            }
            out.push_back(v);
            if (comma == std::string::npos) break;
            start = comma + 1;
        }
        return out;
    }

    template <typename T, int N>
    std::array<T, N> mfind(const std::string &name) const
    {
        const std::vector<T> v = mfind<T>(name);
        if ((int)v.size() != N) {
            std::cerr << "# error: mfind<" << N << ", T>(\"" << name << "\").size() != " << N << std::endl;
            std::exit(1);
        }
        std::array<T, N> out;
        for (int k = 0; k < N; ++k) out[k] = v[k];
        return out;
    }

    template <typename Stream>
    void print(Stream &os) const
    {
        os << "#===== updated arguments =====" << std::endl;
        for (size_t k = 0; k < names_.size(); ++k) os << "# " << std::setw(8) << names_[k] << " : " << values_[k] << std::endl;
        os << "#=============================" << std::endl;
    }

  private:
    const std::string &value(const std::string &name) const
    {
        for (size_t k = 0; k < names_.size(); ++k)
            if (names_[k] == name) return values_[k];
        std::cerr << "# error(out) ---> Threre is no option for your calling. : " << name << std::endl;
        std::exit(1);
    }
    template <typename T>
    static bool cast(const std::string &text, T &v)
    {
        std::stringstream ss(text);
        ss >> v;
        return !ss.fail();
    }
    [[noreturn]] static void help(const char *prog, const std::vector<pair_t> &options,
                                  const std::vector<pair_t> &defaults)
    {
        std::cout << " # option list" << std::endl;
        for (const pair_t &o : options) {
            std::cout << std::setw(8) << o.first << " : " << o.second;
            for (const pair_t &d : defaults)
                if (d.first == o.first) std::cout << " (default : " << d.second << ")";
            std::cout << std::endl;
        }
        std::cout << std::endl << " (hint) " << prog << " -option1=value1 -option2=value2 ..." << std::endl;
        std::exit(1);
    }
    std::vector<std::string> names_, values_;
};

// randNum_generator.cpp: appends nsam rows of L uniform(-1, 1) numbers (precision 10, each
// followed by a tab) to `filename`, writing the header line first when the file is new.  The
// reference seeds std::mt19937 from std::random_device; `seed` makes a file reproducible.
inline void randnum_file_write(const std::string &filename, int L, int nsam, unsigned seed)
{
    const bool exists = std::ifstream(filename).is_open();
    std::ofstream out(filename, exists ? std::ios::app : std::ios::out);
    if (!exists) out << "#       X_1        X_2     ....    " << std::endl;
    out.precision(10);
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (int s = 0; s < nsam; ++s) {
        for (int i = 0; i < L; ++i) out << dist(gen) << "\t";
        out << std::endl;
    }
}

// The drivers' replay of a randNum file (ConservedTotalSpin_Kagome_randomHop.cu:85-107): every
// sample reads the NEXT LINE and extracts n numbers with operator>>.  The reference does not
// skip the header, so the first sample reads "#  X_1 ..." and every extraction fails, giving 0
// (a disorder-free sample); samples past the end of the file also give zeros.  open() reports
// whether the file exists (the reference prints "There is not the file!" and reads zeros).
class RandNumReplay {
  public:
    bool open(const std::string &filename)
    {
        in_.open(filename);
        return in_.is_open();
    }
    void next(int n, std::vector<double> &x)
    {
        std::string line;
        std::getline(in_, line);
        std::istringstream iss(line);
        x.assign((size_t)n, 0.0);
        for (int i = 0; i < n; ++i) {
            double v = 0.0;
            iss >> v; // a failed extraction stores 0 and keeps failing
            x[(size_t)i] = v;
        }
    }

  private:
    std::ifstream in_;
};

// The .dat outputs of the drivers: opened in append mode when the file exists, otherwise
// created with the header line; precision 10; row() writes "\t" before every value and ends the
// line (e.g. ConservedTotalSpin_Kagome_randomHop.cu:67-80, 121-125).
class DatFile {
  public:
    DatFile(const std::string &filename, const std::string &header)
    {
        const bool exists = std::ifstream(filename).is_open();
        out_.open(filename, exists ? std::ios::app : std::ios::out);
        if (!exists) out_ << header << std::endl;
        out_.precision(10);
    }
    bool good() const { return out_.good(); }
    void row(const std::vector<double> &values)
    {
        for (double v : values) out_ << "\t" << v;
        out_ << std::endl;
    }

  private:
    std::ofstream out_;
};

} // namespace ed
