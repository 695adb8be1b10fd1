// Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// engine's host producers and preparation (host_prep.cpp, spectral.cpp,
// adjoint_host.cpp) and the oracle (oracle/tcl4_oracle.cpp), built with g++
// and the sanitizers (no device code), each entry exercised on ordinary,
// edge and error inputs.  Exit status 0 and "host_sanitize: ok" when every
// check holds; a sanitizer report aborts the run.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../taco_amd/csrc/engine_internal.hpp"

using cd = std::complex<double>;
using namespace taco_engine;

// The engine library keeps the last error message with its device code; the
// host sources alone only need somewhere to put it.
namespace taco_engine {
static std::string g_msg;
void set_last_error(const std::string& msg) { g_msg = msg; }
}  // namespace taco_engine

extern "C" {
int orc_bcf_fft(size_t N, double dt, int kind, double p0, double p1, double beta, double* C_out);
int orc_gamma_trapz(const double* C, size_t N, double dt, const double* omegas, int M, double* G);
int orc_gamma_rect(const double* C, size_t N, double dt, const double* omegas, int M, double* G);
int orc_triple_kernels(const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror, double dt,
                       int method, size_t pad, size_t keep, double* F, double* C, double* R, int nthreads);
int orc_l4_series(const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror, const int* p2f,
                  int N, const double* ops, int num_ops, double dt, int method, size_t pad, const size_t* tidx,
                  size_t ntau, double* L4_out, int nthreads);
int orc_tcl4_propagate(const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                       const int* p2f, int N, const double* eps, const double* ops, int num_ops, double dt,
                       int method, size_t pad, size_t steps, int order, double* r_inout, double* samples_out,
                       int nthreads);
int orc_system_build(const double* H, int N, const double* A_lab, int num_ops, double freq_tol, double* eps,
                     double* U, double* A_eig, int* nf_out, double* omegas, int* mirror, int* p2f, int* zero_index);
}

static int failures = 0;
static void check(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

static std::vector<cd> random_hermitian(int n, std::mt19937_64& g) {
    std::normal_distribution<double> nd;
    std::vector<cd> H(static_cast<std::size_t>(n) * n);
    for (int r = 0; r < n; ++r)
        for (int c = r; c < n; ++c) {
            const cd v = r == c ? cd(nd(g), 0.0) : cd(nd(g), nd(g));
            H[r + c * n] = v;
            H[c + r * n] = std::conj(v);
        }
    return H;
}

int main() {
    std::mt19937_64 g(7);
    // system preparation: a random Hermitian H, one with a degenerate pair, a 1 x 1 H
    for (int n : {1, 2, 5, 9}) {
        std::vector<cd> H = random_hermitian(n, g), A = random_hermitian(n, g);
        if (n >= 5) H[1 + 1 * n] = H[0];  // diagonal entries equal
        const std::size_t NN = static_cast<std::size_t>(n) * n;
        std::vector<double> eps(n), om(NN);
        std::vector<cd> U(NN), Ae(NN);
        std::vector<int> mir(NN), p2f(NN);
        int nf = 0, zi = -1;
        system_build(reinterpret_cast<const double*>(H.data()), n, reinterpret_cast<const double*>(A.data()), 1, 1e-9,
                     eps.data(), reinterpret_cast<double*>(U.data()), reinterpret_cast<double*>(Ae.data()), &nf,
                     om.data(), mir.data(), p2f.data(), &zi);
        double res = 0.0;
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) {
                cd s = 0.0;
                for (int m = 0; m < n; ++m) s += H[r + m * n] * U[m + c * n];
                res = std::max(res, std::abs(s - eps[c] * U[r + c * n]));
            }
        check(res < 1e-11, "eigen residual");
        check(nf >= 1 && nf <= static_cast<int>(NN) && zi >= 0 && zi < nf, "bucket count and zero bucket");
        std::vector<double> e2(n), om2(NN);
        std::vector<cd> U2(NN), Ae2(NN);
        std::vector<int> mir2(NN), p2f2(NN);
        int nf2 = 0, zi2 = -1;
        check(orc_system_build(reinterpret_cast<const double*>(H.data()), n,
                               reinterpret_cast<const double*>(A.data()), 1, 1e-9, e2.data(),
                               reinterpret_cast<double*>(U2.data()), reinterpret_cast<double*>(Ae2.data()), &nf2,
                               om2.data(), mir2.data(), p2f2.data(), &zi2) == 0 && nf2 == nf,
              "oracle system build");
    }
    bool threw = false;
    try {
        std::vector<cd> bad = {cd(0, 0), cd(1, 0), cd(0, 0), cd(0, 0)};
        std::vector<double> e(2), om(4);
        std::vector<cd> U(4), Ae(4);
        std::vector<int> m(4), p(4);
        int nf, zi;
        system_build(reinterpret_cast<const double*>(bad.data()), 2, reinterpret_cast<const double*>(bad.data()), 1,
                     1e-9, e.data(), reinterpret_cast<double*>(U.data()), reinterpret_cast<double*>(Ae.data()), &nf,
                     om.data(), m.data(), p.data(), &zi);
    } catch (const Error&) {
        threw = true;
    }
    check(threw, "non-Hermitian H rejected");

    // bath correlation and Gamma: both spectral densities, finite and zero temperature, small N
    for (std::size_t N : {0UL, 1UL, 2UL, 3UL, 100UL, 1000UL}) {
        for (int kind : {0, 1})
            for (double beta : {0.5, static_cast<double>(INFINITY)}) {
                std::vector<cd> C(N + 1), Co(N + 1);
                bcf_fft(N, 0.01, kind, 0.1, 5.0, beta, reinterpret_cast<double*>(C.data()));
                check(orc_bcf_fft(N, 0.01, kind, 0.1, 5.0, beta, reinterpret_cast<double*>(Co.data())) == 0,
                      "oracle bcf");
                const double om[3] = {-2.0, 0.0, 1.5};
                std::vector<cd> G((N + 1) * 3), Gr((N + 1) * 3), Go((N + 1) * 3);
                gamma_trapz(reinterpret_cast<const double*>(C.data()), N + 1, 0.01, om, 3,
                            reinterpret_cast<double*>(G.data()));
                gamma_rect(reinterpret_cast<const double*>(C.data()), N + 1, 0.01, om, 3,
                           reinterpret_cast<double*>(Gr.data()));
                check(orc_gamma_rect(reinterpret_cast<const double*>(C.data()), N + 1, 0.01, om, 3,
                                     reinterpret_cast<double*>(Go.data())) == 0 &&
                          Go == Gr,
                      "rect Gamma equals the oracle");
                check(orc_gamma_trapz(reinterpret_cast<const double*>(C.data()), N + 1, 0.01, om, 3,
                                      reinterpret_cast<double*>(Go.data())) == 0,
                      "oracle trapz");
            }
    }
    threw = false;
    try {
        std::vector<double> C(4);
        bcf_fft(1, -1.0, 0, 1.0, 1.0, 1.0, C.data());
    } catch (const Error&) {
        threw = true;
    }
    check(threw, "dt <= 0 rejected");

    // spectral expressions: evaluation, bath correlation, every error path
    {
        const char* names[2] = {"alpha", "omega_c"};
        const double vals[2] = {0.05, 5.0};
        taco_spectral* J = nullptr;
        check(taco_spectral_create("alpha*w*exp(-w/omega_c)*step(w) + max(0, -w^2) + pow(w, 0.5)*0", names, vals, 2,
                                   &J) == TACO_OK,
              "expression compiles");
        std::vector<double> w(500), out(500);
        for (int k = 0; k < 500; ++k) w[k] = 0.01 * k;
        check(taco_spectral_eval(J, w.data(), w.size(), out.data()) == TACO_OK, "expression evaluates");
        check(std::abs(out[100] - 0.05 * 1.0 * std::exp(-1.0 / 5.0)) < 1e-15, "expression value");
        std::vector<double> Cs(2 * 301);
        check(taco_bcf_fft_spectral(300, 0.01, J, 1.0, Cs.data()) == TACO_OK, "expression bath correlation");
        taco_spectral_destroy(J);
        const char* bad[] = {"w*q", "foo(w)", "(w", "w)", "pow(w)", "exp(w,1)", "2^-1", "w 2", "", "w # 1", "w, 1",
                             "max(1,2,3)", "((((((((((w))))))))))+"};
        for (const char* b : bad) {
            taco_spectral* Jb = nullptr;
            check(taco_spectral_create(b, nullptr, nullptr, 0, &Jb) == TACO_ERR_INVALID && Jb == nullptr,
                  "bad expression rejected");
        }
        // a long program (deeper than the evaluator's fixed stack)
        std::string longe = "w";
        for (int k = 0; k < 200; ++k) longe = "(" + longe + "+w*" + std::to_string(k) + ")";
        check(taco_spectral_create(longe.c_str(), nullptr, nullptr, 0, &J) == TACO_OK, "long expression");
        double one = 1.0, v = 0.0;
        check(taco_spectral_eval(J, &one, 1, &v) == TACO_OK && v == 1.0 + 19900.0, "long expression value");
        taco_spectral_destroy(J);
    }

    // phase tables, correctly rounded routines, twiddles, FFT sizes
    {
        const double omg[3] = {-1.25, 0.0, 3.5};
        std::vector<double> tab(2 * 3 * 257);
        libm_phase_table(omg, 3, 257, 0.01, tab.data());
        check(tab[0] == 1.0 && tab[1] == 0.0, "phase table row start");
        std::vector<double> x(1000), s(1000), c(1000), e(1000);
        for (int k = 0; k < 1000; ++k) x[k] = -50.0 + 0.1 * k;
        cr_sincos_host(x.data(), x.size(), s.data(), c.data());
        cr_exp_host(x.data(), x.size(), e.data());
        double err = 0.0;
        for (int k = 0; k < 1000; ++k)
            err = std::max({err, std::abs(s[k] - std::sin(x[k])), std::abs(c[k] - std::cos(x[k])),
                            std::abs(e[k] - std::exp(x[k])) / std::exp(x[k])});
        check(err < 1e-15, "correctly rounded routines");
        for (int L : {1, 2, 5, 12}) {
            std::vector<double> t1, t2;
            make_twiddles(L, t1);
            make_twiddles_exact(L, t2);
            check(t1.size() == t2.size() && !t1.empty(), "twiddle tables");
        }
        check(next_pow2(1000) == 1024 && fcr_nfft(1000, 0) >= 1999, "FFT sizes");
    }

    // the parallel stable sort used to group triples equals std::stable_sort
    for (long n : {1000L, 300000L}) {
        std::vector<std::uint64_t> key(static_cast<std::size_t>(n));
        for (auto& k : key) k = g() % 4000;
        std::vector<int> o1(static_cast<std::size_t>(n));
        for (long q = 0; q < n; ++q) o1[static_cast<std::size_t>(q)] = static_cast<int>(n - 1 - q);
        std::vector<int> o2 = o1;
        std::stable_sort(o1.begin(), o1.end(), [&](int x, int y) { return key[x] < key[y]; });
        stable_sort_by_key(o2, key.data());
        check(o1 == o2, "parallel stable sort");
    }

    // TCL2 reverse helpers: parts and operator cotangents on a d = 3 system
    {
        const int n = 3, nf = 7;
        std::vector<cd> A = random_hermitian(n, g);
        std::vector<int> p2f(9);
        for (int e = 0; e < 9; ++e) p2f[e] = e % nf;
        Tcl2Parts parts = tcl2_parts(A.data(), 1, n, p2f.data(), nf);
        check(!parts.b.empty() && parts.P.size() == parts.b.size() * 9, "TCL2 parts");
        std::vector<cd> W(static_cast<std::size_t>(nf) * 81), V(static_cast<std::size_t>(nf) * 9), ob(9);
        std::normal_distribution<double> nd;
        for (auto& z : W) z = cd(nd(g), nd(g));
        for (auto& z : V) z = cd(nd(g), nd(g));
        tcl2_ops_adjoint(parts, W.data(), V.data(), n, p2f.data(), ob.data());
        check(std::isfinite(std::abs(ob[4])), "TCL2 operator cotangents");
    }

    // the oracle's F/C/R (both methods), L4 series and propagation on a small system
    {
        const int n = 2;
        std::vector<cd> H = random_hermitian(n, g), A = random_hermitian(n, g);
        std::vector<double> eps(n), om(4);
        std::vector<cd> U(4), Ae(4);
        std::vector<int> mir(4), p2f(4);
        int nf = 0, zi = 0;
        orc_system_build(reinterpret_cast<const double*>(H.data()), n, reinterpret_cast<const double*>(A.data()), 1,
                         1e-9, eps.data(), reinterpret_cast<double*>(U.data()), reinterpret_cast<double*>(Ae.data()),
                         &nf, om.data(), mir.data(), p2f.data(), &zi);
        const std::size_t Nt = 64;
        std::vector<cd> C(Nt), G(Nt * nf);
        orc_bcf_fft(Nt - 1, 0.02, 0, 0.1, 5.0, 1.0, reinterpret_cast<double*>(C.data()));
        orc_gamma_trapz(reinterpret_cast<const double*>(C.data()), Nt, 0.02, om.data(), nf,
                        reinterpret_cast<double*>(G.data()));
        const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
        for (int method : {0, 1}) {
            std::vector<cd> F(nf3 * Nt), Cc(nf3 * Nt), R(nf3 * Nt);
            check(orc_triple_kernels(reinterpret_cast<const double*>(G.data()), Nt, nf, om.data(), mir.data(), 0.02,
                                     method, 0, Nt, reinterpret_cast<double*>(F.data()),
                                     reinterpret_cast<double*>(Cc.data()), reinterpret_cast<double*>(R.data()),
                                     2) == 0,
                  "oracle F/C/R");
            const size_t tidx[3] = {0, 5, 63};
            std::vector<cd> L4(3 * 16);
            check(orc_l4_series(reinterpret_cast<const double*>(G.data()), Nt, nf, om.data(), mir.data(),
                                p2f.data(), n, reinterpret_cast<const double*>(Ae.data()), 1, 0.02, method, 0, tidx,
                                3, reinterpret_cast<double*>(L4.data()), 2) == 0,
                  "oracle L4 series");
            std::vector<cd> r = {cd(1, 0), cd(0, 0), cd(0, 0), cd(0, 0)}, samples(31 * 4);
            check(orc_tcl4_propagate(reinterpret_cast<const double*>(G.data()), Nt, nf, om.data(), mir.data(),
                                     p2f.data(), n, eps.data(), reinterpret_cast<const double*>(Ae.data()), 1, 0.02,
                                     method, 0, 30, 4, reinterpret_cast<double*>(r.data()),
                                     reinterpret_cast<double*>(samples.data()), 2) == 0,
                  "oracle propagation");
            check(std::abs((samples[30 * 4] + samples[30 * 4 + 3]) - 1.0) < 1e-10, "trace preserved");
        }
    }
    if (failures) {
        std::printf("host_sanitize: %d failure(s)\n", failures);
        return 1;
    }
    std::printf("host_sanitize: ok\n");
    return 0;
}
