"""Driver I/O of the reference's programs, for Python drivers: the randNum disorder files and
the append-mode .dat outputs.  Mirrors include/ed/driver_io.hpp (the C++ drivers' version), byte
for byte:

* randnum_write: TotalSpin/src/randNum_generator.cpp -- header line, then rows of L
  uniform(-1, 1) numbers from std::mt19937 (seeded here; the reference seeds from
  std::random_device) through libstdc++'s uniform_real_distribution (generate_canonical with two
  32-bit draws), printed with 10 significant digits, each followed by a tab.
* RandNumReplay: the drivers' replay (ConservedTotalSpin_Kagome_randomHop.cu:85-107) -- one
  line per sample, numbers read with operator>>; the header line is read as the first sample and
  gives zeros, as in the reference, and so do samples past the end of the file.
* DatFile: append mode when the file exists, otherwise created with the header line; each row
  is tab-prefixed values with 10 significant digits.

Command-line parsing stays with Python's argparse; the C++ drivers use ed::argsparse.
"""
from __future__ import annotations

import os

import numpy as np

RANDNUM_HEADER = "#       X_1        X_2     ....    "


def _fmt(v: float) -> str:
    """std::ostream << double with precision 10 (default float field)"""
    return format(float(v), ".10g")


def mt19937_uniform(seed: int, n: int) -> np.ndarray:
    """n draws of std::uniform_real_distribution<double>(-1, 1) on std::mt19937(seed) as
    libstdc++ computes them: u = (g1 + g2 * 2^32) / 2^64 from two 32-bit outputs (the nextafter
    clamp of generate_canonical applies only at u = 1), x = (b - a) u + a."""
    rs = np.random.RandomState(seed)  # init_genrand(seed): the std::mt19937(seed) state
    g = rs.randint(0, 2 ** 32, size=2 * n, dtype=np.uint64).astype(np.float64)
    u = (g[0::2] + g[1::2] * 4294967296.0) / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return u * 2.0 + -1.0


def randnum_write(filename: str, L: int, nsam: int, seed: int) -> None:
    new = not os.path.exists(filename)
    x = mt19937_uniform(seed, L * nsam).reshape(nsam, L)
    with open(filename, "w" if new else "a") as f:
        if new:
            f.write(RANDNUM_HEADER + "\n")
        for row in x:
            f.write("".join(_fmt(v) + "\t" for v in row) + "\n")


class RandNumReplay:
    def __init__(self, filename: str):
        self.exists = os.path.exists(filename)
        self._lines = open(filename).read().split("\n") if self.exists else []
        self._k = 0

    def next(self, n: int) -> np.ndarray:
        line = self._lines[self._k] if self._k < len(self._lines) else ""
        self._k += 1
        out = np.zeros(n)
        for i, tok in enumerate(line.split()[:n]):
            try:
                out[i] = float(tok)
            except ValueError:  # operator>> fails here and for every later number
                break
        return out


class DatFile:
    def __init__(self, filename: str, header: str):
        new = not os.path.exists(filename)
        self._f = open(filename, "w" if new else "a")
        if new:
            self._f.write(header + "\n")
            self._f.flush()

    def row(self, values) -> None:
        self._f.write("".join("\t" + _fmt(v) for v in values) + "\n")
        self._f.flush()

    def close(self) -> None:
        self._f.close()
