// engine_internal.hpp — shared declarations of the TCL4 engine sources.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/taco_tcl4.h"

namespace taco_engine {

// Exception carrying a C-ABI status code; converted at the ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

std::size_t next_pow2(std::size_t n);
std::size_t fcr_nfft(std::size_t Nt, std::size_t pad);
void make_twiddles(int L, std::vector<double>& out);
// The same table with every entry exp(-2 pi i k / 2^s) rounded once from a
// quad-precision evaluation (the window mode's accurate twiddles).
void make_twiddles_exact(int L, std::vector<double>& out);
// F/C/R FFT mode (process-global): TACO_FFT_PARITY or TACO_FFT_WINDOW.
int fft_mode();
void set_fft_mode(int mode);
void bcf_fft(std::size_t N, double dt, int kind, double p0, double p1, double beta, double* C_out);
// C(t_n), n = 0..N, of correlation_fft.hpp:154-214 for any spectral density J
// (called at the sampled frequencies w > 0 only).
void bcf_fft_fn(std::size_t N, double dt, const std::function<double(double)>& J, double beta, double* C_out);
// Rectangle-rule Gamma (gamma.hpp:127-160), columns of N rows per omega.
void gamma_rect(const double* C, std::size_t N, double dt, const double* omegas, int M, double* G);
void gamma_trapz(const double* C, std::size_t N, double dt, const double* omegas, int M, double* G);
void mirror_indices(const double* omegas, int nf, double fidx_tol, int* mirror);
// std::stable_sort(order, key[x] < key[y]) on all host cores (same result).
void stable_sort_by_key(std::vector<int>& order, const std::uint64_t* key);
// key[t] = the bit pattern of (w_i + w_j) + w_k for t = (i nf + j) nf + k (OpenMP).
void triple_omega_keys(const double* omegas, int nf, std::uint64_t* key);
void system_build(const double* H, int N, const double* A_lab, int num_ops, double freq_tol, double* eps,
                  double* U, double* A_eig, int* nf_out, double* omegas, int* mirror, int* p2f,
                  int* zero_index);

// Reverse of the TCL2 generator, host side (adjoint_host.cpp).  Parts: one
// per (operator, bucket) with a nonzero masked operator P: AdA = P^H P
// [N^2], P [N^2]; complex values column-major.  The device forms each part's
// K = dL/dgam [D^2] from P and AdA (k_tcl2_part_k).
struct Tcl2Parts {
    std::vector<std::complex<double>> AdA, P;
    std::vector<int> al, b;
};
Tcl2Parts tcl2_parts(const std::complex<double>* ops, int num_ops, int N, const int* p2f, int nf);
// Adds the operator cotangents formed from the time sums W [nf][D^2] and
// V [nf][N^2] of the device reverse to ops_bar [num_ops][N^2].
void tcl2_ops_adjoint(const Tcl2Parts& parts, const std::complex<double>* W, const std::complex<double>* V, int N,
                      const int* p2f, std::complex<double>* ops_bar);

// Phase and exponential convention (TACO_PHASE_LIBM or TACO_PHASE_CR),
// process-global like the pad factor.
int phase_mode();
void set_phase_mode(int mode);
// exp(x) in the current convention: the platform libm's std::exp, or the
// correctly rounded value.
double exp_in_mode(double x);
// exp(i x) = (cos x, sin x) in the current convention.
std::complex<double> expi_in_mode(double x);
// out[r][k] = std::exp(i * (omega[r] * (k * dt))), k < len, interleaved
// (re, im): the platform libm's values of the reference's F/C/R phases,
// evaluated on all host cores.
void libm_phase_table(const double* omega, std::size_t nrows, std::size_t len, double dt, double* out);

void set_last_error(const std::string& msg);
const double* cr_table();
void cr_sincos_host(const double* x, std::size_t n, double* s, double* c);
const double* exp_table();
void cr_exp_host(const double* x, std::size_t n, double* out);

template <class Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        set_last_error("");
        return TACO_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("out of host memory");
        return TACO_ERR_RUNTIME;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return TACO_ERR_RUNTIME;
    }
}

}  // namespace taco_engine
