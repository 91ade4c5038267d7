// ed/ed.hpp -- C++ interface of the MI355X exact-diagonalization library, with the class
// and method names, argument meaning and call order of the reference's headers, so a driver
// written against them (MBL_test/include/fermi_operator.h, HubMatrix_GPU.cuh,
// HubMatrix_TotalSpin.cuh, SymEigsSolver.cuh, entangle_entropy_GPU.cuh) compiles against this
// header with its include lines changed and device_vector taken from namespace ed.
//
// Header-only; every computation goes through the C ABI of libedhip.so (include/edhip.h).
// Errors: the reference prints (CUDAERR) or asserts; here every failing call throws ed::Error
// carrying ed_last_error().  Vectors live on the device (ed::device_vector, hipMalloc'd).
//
// Build a driver:  hipcc -std=c++17 -I<repo>/include driver.cpp -L<pkg> -ledhip
//                  -Wl,-rpath,<pkg>      (<pkg> = the directory holding libedhip.so)
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../edhip.h"

namespace ed {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(int status, const char *what)
{
    if (status != ED_OK)
        throw Error(status, std::string(what) + ": " + (ed_last_error() ? ed_last_error() : "") + " (status " +
                                std::to_string(status) + ")");
}

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Error(ED_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct is_cplx : std::false_type {};
template <>
struct is_cplx<std::complex<double>> : std::true_type {};

// thrust::device_vector<T> stand-in: contiguous device storage, element type double or
// std::complex<double> (same layout as the C ABI's interleaved complex128)
template <typename T>
class device_vector {
  public:
    device_vector() = default;
    explicit device_vector(size_t n, T value = T()) { resize(n, value); }
    explicit device_vector(const std::vector<T> &h) { assign(h); }
    device_vector(const device_vector &o) { *this = o; }
    device_vector(device_vector &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    device_vector &operator=(const device_vector &o)
    {
        if (this != &o) {
            reset(o.n_);
            if (n_) hip_check(hipMemcpy(p_, o.p_, bytes(), hipMemcpyDeviceToDevice), "device_vector copy");
        }
        return *this;
    }
    device_vector &operator=(device_vector &&o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~device_vector() { (void)hipFree(p_); }

    size_t size() const { return n_; }
    T *data() { return p_; }
    const T *data() const { return p_; }
    void resize(size_t n, T value = T())
    {
        reset(n);
        if (n_) assign(std::vector<T>(n_, value));
    }
    void clear() { reset(0); }
    void assign(const std::vector<T> &h)
    {
        if (h.size() != n_) reset(h.size());
        if (n_) hip_check(hipMemcpy(p_, h.data(), bytes(), hipMemcpyHostToDevice), "device_vector upload");
    }
    std::vector<T> to_host() const
    {
        std::vector<T> h(n_);
        if (n_) hip_check(hipMemcpy(h.data(), p_, bytes(), hipMemcpyDeviceToHost), "device_vector download");
        return h;
    }

  private:
    size_t bytes() const { return n_ * sizeof(T); }
    void reset(size_t n)
    {
        if (n == n_) return;
        (void)hipFree(p_);
        p_ = nullptr;
        n_ = n;
        if (n_) hip_check(hipMalloc((void **)&p_, bytes()), "device_vector alloc");
    }
    T *p_ = nullptr;
    size_t n_ = 0;
};

} // namespace ed

#include "fermi_operator.hpp"
#include "hubmatrix.hpp"
#include "totalspin.hpp"
#include "spectra.hpp"
#include "entangle.hpp"
