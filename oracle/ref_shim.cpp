// ref_shim.cpp -- exposes the reference's header-only kernels through a C ABI
// so tests can check the oracle against them.  TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile against the headers where they lie under the
// reference tree (include/fringe/ulongmask.hpp, include/fringe/EigenLapack.hpp,
// src/nmap/KS2sample.hpp, src/nmap/AD2unique.hpp) and linked with the LAPACK
// that provides zheevr_/zpotrf_/zpotri_ (libmkl_rt here).  Output goes to
// oracle/_ref/ only.  None of the reference's own sources are copied.
#include <complex>
#include <cstring>
#include <vector>
#include <cmath>
#include "fringe/ulongmask.hpp"
#include "fringe/EigenLapack.hpp"
#include "KS2sample.hpp"
#include "AD2unique.hpp"

// The reference KS walk reads one element past each array (KS2sample.hpp:120,
// :125); callers' arrays are copied into buffers with a NaN sentinel so that
// read stays in bounds and never compares equal.
extern "C" double ref_ks2_test(const float* a, const float* b, int n) {
    std::vector<float> aa(a, a + n), bb(b, b + n);
    aa.push_back(NAN);
    bb.push_back(NAN);
    KS2sample w(n);
    return w.test(aa.data(), bb.data());
}

extern "C" double ref_kolmogorov_prob(double z) {
    KS2sample w(1);
    return w.KolmogorovProb(z);
}

extern "C" double ref_ad2_test(const float* a, const float* b, int n) {
    AD2unique w(n);
    return w.test(a, b);
}

extern "C" void ref_setbit(unsigned int* arr, int Ny, int Nx, int ii, int jj) {
    const Ulongmask m = {Ny, Nx};
    m.setbit(arr, ii, jj, true);
}

extern "C" int ref_getbit(unsigned int* arr, int Ny, int Nx, int ii, int jj) {
    const Ulongmask m = {Ny, Nx};
    return m.getbit(arr, ii, jj) ? 1 : 0;
}

// A: column-major n x n complex<double> as interleaved doubles; preserved.
extern "C" int ref_eigen_index(const double* A, int n, int index, int want_vec,
                               double* eigval, double* eigvec) {
    EVWorker w(n);
    std::vector<std::complex<double>> a((size_t)n * n);
    std::memcpy(a.data(), A, sizeof(double) * 2 * n * n);
    int info = w.eigenWithIndex(a.data(), index, want_vec != 0);
    *eigval = w.eigval[0];
    if (want_vec) std::memcpy(eigvec, w.eigvec, sizeof(double) * 2 * n);
    return info;
}

extern "C" int ref_pd_inverse(double* A, int n) {
    EVWorker w(n);
    return w.positiveDefiniteInverse(reinterpret_cast<std::complex<double>*>(A));
}
