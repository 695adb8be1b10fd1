// TEST INFRASTRUCTURE ONLY — the CPU oracle for the TCL4 hot path.
//
// This file is a C++17 restatement of the reference serial/OpenMP algorithm
// (jcarrots/taco, cpp/src/tcl + cpp/include/taco), written without Eigen.
// It keeps the reference's operation order wherever rounding depends on it:
// the radix-2 FFT with bit reversal and recurrence twiddles, the Nfft rule,
// `(sum y)*dt` prefix integrals, `std::exp` phases of `omega*(k*dt)`, and
// plain (non-FMA) complex products.  Build flags pin `-ffp-contract=off`.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker.  Nothing in taco_amd/ links it.
//
// Parity status of each function is listed in the header of oracle/oracle.py.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include <quadmath.h>

namespace {

using cd = std::complex<double>;
constexpr double PI = 3.141592653589793238462643383279502884;

thread_local std::string g_err;

// ---------------------------------------------------------------------------
// FFT — correlation_fft.hpp:27-73 (next_pow2, fft_inplace_builtin_dir)
// ---------------------------------------------------------------------------
std::size_t next_pow2(std::size_t n) {
    if (n == 0) return 1;
    --n;
    n |= n >> 1;  n |= n >> 2;  n |= n >> 4;
    n |= n >> 8;  n |= n >> 16; n |= n >> 32;
    return ++n;
}

// In-place radix-2 DIT with bit reversal; twiddles by the recurrence w *= wlen,
// restarted at 1 for every block; the inverse scales by 1/n at the end.
void fft_ref(cd* a, std::size_t n, bool forward) {
    if (n == 0) return;
    if ((n & (n - 1)) != 0) throw std::runtime_error("fft: length must be power of two");
    std::size_t j = 0;
    for (std::size_t i = 1; i < n; ++i) {
        std::size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (std::size_t len = 2; len <= n; len <<= 1) {
        const double ang = (forward ? -1.0 : +1.0) * 2.0 * PI / static_cast<double>(len);
        const cd wlen(std::cos(ang), std::sin(ang));
        const std::size_t half = len >> 1;
        for (std::size_t i = 0; i < n; i += len) {
            cd w(1.0, 0.0);
            for (std::size_t k = 0; k < half; ++k) {
                const cd u = a[i + k];
                const cd v = w * a[i + k + half];
                a[i + k] = u + v;
                a[i + k + half] = u - v;
                w *= wlen;
            }
        }
    }
    if (!forward) {
        const double inv_n = 1.0 / static_cast<double>(n);
        for (std::size_t i = 0; i < n; ++i) a[i] *= inv_n;
    }
}

// Nfft rule shared by causal_conv_fft and the fused GPU builders:
// tcl4_kernels.cpp:40-49 — next_pow2(max(2N-1, pad*N)), at least 2.
std::size_t fcr_nfft(std::size_t N, std::size_t pad) {
    const std::size_t L = 2 * N - 1;
    std::size_t target = L;
    if (pad > 0) {
        target = pad * N;
        if (target < L) target = L;
    }
    std::size_t n = next_pow2(target);
    if (n < 2) n = 2;
    return n;
}

// ---------------------------------------------------------------------------
// Spectral densities used by the reference drivers and configs.
// kind 0: Ohmic  J = a*w*exp(-w/wc)   (TCL4_spin_boson_example.cpp:210-213)
// kind 1: Drude–Lorentz J = 2*lambda*gamma*w/(w^2+gamma^2)
// Both vanish for w <= 0, as the reference lambdas do.
// ---------------------------------------------------------------------------
// Exponential of the bath correlation: std::exp as the reference calls it
// (phase mode 0) or the correctly rounded value (mode 1, the engine's
// convention): exp in quad precision rounded once to double (the quad result
// is within ~2^-112 of exp(x), so the rounding is correct unless exp(x) lies
// that close to a rounding boundary; also correct in the subnormal range).
extern int g_phase_mode;
cd phase(double arg);
double bath_exp(double x) {
    if (g_phase_mode == 0) return std::exp(x);
    return static_cast<double>(expq(static_cast<__float128>(x)));
}

double spectral_J(int kind, double p0, double p1, double w) {
    if (!(w > 0.0)) return 0.0;
    if (kind == 0) return p0 * w * bath_exp(-w / p1);
    return 2.0 * p0 * p1 * w / (w * w + p1 * p1);
}

// bcf_fft_fun — correlation_fft.hpp:154-214 (forward FFT, scale dω/π).
void bcf_fft(std::size_t N, double dt, int kind, double p0, double p1, double beta, cd* C_out) {
    if (!(dt > 0.0)) throw std::invalid_argument("dt must be > 0");
    const std::size_t Nt = N + 1;
    std::size_t Nfft = 2 * Nt;
    if ((Nfft & (Nfft - 1)) != 0) Nfft = next_pow2(Nfft);
    const double tf = dt * static_cast<double>(Nfft);
    const double domega = 2.0 * PI / tf;
    const double omegaNy = PI / dt;
    const std::size_t posCount = Nfft / 2 - 1;
    std::vector<cd> S(Nfft, cd(0.0, 0.0));
    if (std::isinf(beta)) {
        for (std::size_t k = 1; k <= posCount; ++k) {
            const double w = domega * static_cast<double>(k);
            S[k] = cd((PI / 2.0) * spectral_J(kind, p0, p1, w), 0.0);
        }
        S[0] = cd(0.0, 0.0);
        S[Nfft / 2] = cd(0.25 * PI * spectral_J(kind, p0, p1, omegaNy), 0.0);
    } else {
        const double r = bath_exp(-beta * domega);
        double KMS = r;
        for (std::size_t k = 1; k <= posCount; ++k) {
            const double w = domega * static_cast<double>(k);
            const double Jval = spectral_J(kind, p0, p1, w);
            const double Sp = (PI / 2.0) * Jval / (1.0 - KMS);
            S[k] = cd(Sp, 0.0);
            S[Nfft - k] = cd(Sp * KMS, 0.0);
            KMS *= r;
        }
        const double ed = r;
        S[0] = cd(0.5 * (PI / 2.0) * spectral_J(kind, p0, p1, domega) * (1.0 + ed) / (1.0 - ed), 0.0);
        const double eNy = bath_exp(beta * omegaNy);
        S[Nfft / 2] = cd(0.25 * PI * spectral_J(kind, p0, p1, omegaNy) * (1.0 + 2.0 / (eNy - 1.0)), 0.0);
    }
    fft_ref(S.data(), Nfft, true);
    const double scale = domega / PI;
    for (std::size_t n = 0; n < Nt; ++n) C_out[n] = S[n] * scale;
}

// compute_trapz_prefix_multi_matrix — gamma.hpp:166-213 (phase recurrence).
void gamma_trapz(const cd* C, std::size_t N, double dt, const double* omegas, int M, cd* G) {
    if (N == 0 || M == 0 || !(dt > 0.0)) return;
    const cd half_dt(dt / 2.0, 0.0);
    for (int j = 0; j < M; ++j) {
        cd* col = G + static_cast<std::size_t>(j) * N;
        col[0] = cd(0.0, 0.0);
        const cd step = phase(omegas[j] * dt);
        cd phi(1.0, 0.0), acc(0.0, 0.0), prev = C[0];
        for (std::size_t k = 1; k < N; ++k) {
            const cd phi_next = phi * step;
            acc += half_dt * (phi * prev + phi_next * C[k]);
            col[k] = acc;
            phi = phi_next;
            prev = C[k];
        }
    }
}

// compute_rect_prefix_multi_matrix — gamma.hpp:127-160.
void gamma_rect(const cd* C, std::size_t N, double dt, const double* omegas, int M, cd* G) {
    if (N == 0 || M == 0 || !(dt > 0.0)) return;
    for (int j = 0; j < M; ++j) {
        cd* col = G + static_cast<std::size_t>(j) * N;
        const cd step = std::exp(cd{0.0, omegas[j] * dt});
        cd phi(1.0, 0.0);
        cd acc = dt * phi * C[0];
        col[0] = acc;
        for (std::size_t k = 1; k < N; ++k) {
            phi *= step;
            acc += dt * phi * C[k];
            col[k] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// F/C/R series — tcl4_kernels.cpp:19-175 (Convolution) and :188-281 (Direct).
// All functions produce the first `keep` samples of the length-Nt series;
// with the Convolution method the FFT length still follows the full Nt,
// exactly as the reference computes the whole series and a caller slices it
// (tcl4_mpi_omp.cpp:82-156 keeps only the requested t in the same way).
// ---------------------------------------------------------------------------
// Phase convention.  The reference forms exp(+-i*omega*t_k) with std::exp of a
// purely imaginary complex (tcl4_kernels.cpp:116-120), i.e. the platform
// libm's sincos, whose last bits differ between libms (glibc is not
// correctly rounded).  Mode 0 calls std::exp exactly as the reference does on
// this platform; mode 1 uses the correctly rounded cos/sin of the same double
// argument omega*(k*dt), the platform-independent value.
int g_phase_mode = 0;

// Round a long-double approximation (error <= 2 ulp of long double) of
// f(x) to double; when it lies too close to a rounding boundary, decide with
// the quad-precision value instead.
template <class QF>
double round_to_double(long double v, double x, QF qf) {
    const double d = static_cast<double>(v);
    if (v == 0.0L) return d;
    const long double r = v - static_cast<long double>(d);
    const double nb = std::nextafter(d, r >= 0 ? std::numeric_limits<double>::infinity()
                                               : -std::numeric_limits<double>::infinity());
    const long double half = std::fabs(static_cast<long double>(nb) - static_cast<long double>(d)) / 2.0L;
    const long double ulp_ld = std::fabs(v) * 0x1p-63L;
    if (std::fabs(std::fabs(r) - half) <= 16.0L * ulp_ld) return static_cast<double>(qf(static_cast<__float128>(x)));
    return d;
}

void cr_sincos(double x, double* s, double* c) {
    const long double xl = x;
    *s = round_to_double(sinl(xl), x, [](__float128 q) { return sinq(q); });
    *c = round_to_double(cosl(xl), x, [](__float128 q) { return cosq(q); });
}

cd phase(double arg) {
    if (g_phase_mode == 0) return std::exp(cd{0.0, arg});
    double s, c;
    cr_sincos(arg, &s, &c);
    return cd(c, s);
}

struct Phases {
    std::vector<cd> plus, minus;
    void build(std::size_t n, double omega, double dt) {
        plus.resize(n); minus.resize(n);
        for (std::size_t k = 0; k < n; ++k) {
            const double tk = static_cast<double>(k) * dt;
            plus[k] = phase(omega * tk);
            minus[k] = phase(-omega * tk);
        }
    }
};

// prefix_int_left_vec — accum += y_k; out_k = accum * dt.
void prefix_int_left(const cd* y, std::size_t n, double dt, cd* out) {
    cd accum(0.0, 0.0);
    for (std::size_t k = 0; k < n; ++k) {
        accum += y[k];
        out[k] = accum * dt;
    }
}

// Spectrum of a zero-padded series (first half of causal_conv_fft).
void padded_spectrum(const cd* x, std::size_t n, std::size_t Nfft, cd* out) {
    for (std::size_t k = 0; k < Nfft; ++k) out[k] = (k < n) ? x[k] : cd(0.0, 0.0);
    fft_ref(out, Nfft, true);
}

// Second half of causal_conv_fft: out_k = dt * ifft(Fa .* Gb)[k], k < keep.
void conv_from_spectra(const cd* Fa, const cd* Gb, std::size_t Nfft, double dt,
                       std::size_t keep, cd* work, cd* out) {
    for (std::size_t k = 0; k < Nfft; ++k) work[k] = Fa[k] * Gb[k];
    fft_ref(work, Nfft, false);
    for (std::size_t k = 0; k < keep; ++k) out[k] = dt * work[k];
}

// F (op = 0) or C (op = 1) via Convolution, given the spectrum of the g2
// operand (g2 for F, conj(g2) for C) so the per-bucket spectra are shared.
// The values are identical to recomputing them per triple: the same FFT runs
// on the same inputs.  Fa_ws holds the spectrum of g1*phase_minus.
void fc_convolution(const cd* g1, const cd* g2op, const cd* G2hat, const Phases& ph,
                    std::size_t Nt, std::size_t Nfft, double dt, std::size_t keep,
                    const cd* g1_phase, const cd* Fa, cd* work, cd* conv, cd* y, cd* prefix,
                    cd* out) {
    (void)g1;
    for (std::size_t k = 0; k < keep; ++k) y[k] = g2op[k] * ph.plus[k];
    prefix_int_left(y, keep, dt, prefix);
    conv_from_spectra(Fa, G2hat, Nfft, dt, keep, work, conv);
    for (std::size_t k = 0; k < keep; ++k) {
        const cd term1 = g1_phase[k] * prefix[k];
        out[k] = term1 - conv[k];
    }
    (void)Nt;
}

void r_convolution(const cd* g1, const cd* g2, const Phases& ph, std::size_t keep, double dt,
                   cd* y, cd* prefix, cd* prefixP, cd* out) {
    for (std::size_t k = 0; k < keep; ++k) y[k] = g2[k] * ph.minus[k];
    prefix_int_left(y, keep, dt, prefix);
    for (std::size_t k = 0; k < keep; ++k) {
        const cd P = g1[k] * g2[k];
        y[k] = P * ph.minus[k];
    }
    prefix_int_left(y, keep, dt, prefixP);
    for (std::size_t k = 0; k < keep; ++k) {
        const cd term1 = g1[k] * prefix[k];
        out[k] = term1 - prefixP[k];
    }
}

// Direct O(Nt^2) forms, tcl4_kernels.cpp:188-281.
void f_direct(const cd* g1, const cd* g2, const Phases& ph, std::size_t keep, double dt,
              bool conj_g2, cd* out) {
    cd prefix(0.0, 0.0);
    for (std::size_t k = 0; k < keep; ++k) {
        const cd g2k = conj_g2 ? std::conj(g2[k]) : g2[k];
        prefix += g2k * ph.plus[k] * dt;
        const cd first = g1[k] * prefix * ph.minus[k];
        cd second(0.0, 0.0);
        for (std::size_t m = 0; m <= k; ++m) {
            const cd g2m = conj_g2 ? std::conj(g2[m]) : g2[m];
            second += g1[k - m] * g2m * ph.minus[k - m];
        }
        out[k] = first - dt * second;
    }
}

void r_direct(const cd* g1, const cd* g2, const Phases& ph, std::size_t keep, double dt, cd* out) {
    cd p2(0.0, 0.0), pP(0.0, 0.0);
    for (std::size_t k = 0; k < keep; ++k) {
        p2 += g2[k] * ph.minus[k] * dt;
        const cd term1 = g1[k] * p2;
        const cd Pk = g1[k] * g2[k];
        pP += Pk * ph.minus[k] * dt;
        out[k] = term1 - pP;
    }
}

// Mirror index — tcl4.cpp:132-146 (identical to build_map, tcl4.cpp:62-86).
void mirror_indices(const double* omegas, int nf, double fidx_tol, int* mirror) {
    const double tol = std::max(1e-12, fidx_tol);
    for (int j = 0; j < nf; ++j) {
        const double w = omegas[j];
        if (std::abs(w) <= tol) { mirror[j] = j; continue; }
        const double target = -w;
        double best = std::numeric_limits<double>::infinity();
        int best_idx = -1;
        for (int jp = 0; jp < nf; ++jp) {
            const double dw = std::abs(omegas[jp] - target);
            if (dw < best) { best = dw; best_idx = jp; }
        }
        mirror[j] = best_idx >= 0 ? best_idx : j;
    }
}

// compute_triple_kernels — tcl4.cpp:103-183.  Output layout:
// out[(idx3 * keep) + t], idx3 = (i*nf + j)*nf + k.
void triple_kernels(const cd* gamma, std::size_t Nt, int nf, const double* omegas, const int* mirror,
                    double dt, int method, std::size_t pad, std::size_t keep,
                    cd* Fo, cd* Co, cd* Ro, int nthreads) {
    if (keep > Nt) throw std::out_of_range("keep > Nt");
    const std::size_t Nfft = fcr_nfft(Nt, pad);
    const bool conv = (method == 0);
    // Per-bucket spectra of g and conj(g) (the B operands of F and C).
    std::vector<cd> Ghat, Gchat;
    if (conv) {
        Ghat.resize(static_cast<std::size_t>(nf) * Nfft);
        Gchat.resize(static_cast<std::size_t>(nf) * Nfft);
        #pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int b = 0; b < nf; ++b) {
            const cd* g = gamma + static_cast<std::size_t>(b) * Nt;
            padded_spectrum(g, Nt, Nfft, Ghat.data() + static_cast<std::size_t>(b) * Nfft);
            std::vector<cd> gc(Nt);
            for (std::size_t t = 0; t < Nt; ++t) gc[t] = std::conj(g[t]);
            padded_spectrum(gc.data(), Nt, Nfft, Gchat.data() + static_cast<std::size_t>(b) * Nfft);
        }
    }
    const std::ptrdiff_t total = static_cast<std::ptrdiff_t>(nf) * nf;
    #pragma omp parallel num_threads(nthreads)
    {
        Phases ph;
        std::vector<cd> g1_phase(Nt), Fa(conv ? Nfft : 0), work(conv ? Nfft : 0);
        std::vector<cd> convb(keep), y(keep), prefix(keep), prefixP(keep), g2c(keep);
        #pragma omp for schedule(dynamic)
        for (std::ptrdiff_t idx = 0; idx < total; ++idx) {
            const int i = static_cast<int>(idx / nf);
            const int j = static_cast<int>(idx % nf);
            const cd* g1 = gamma + static_cast<std::size_t>(i) * Nt;
            const cd* g2 = gamma + static_cast<std::size_t>(j) * Nt;
            const int jm = mirror[j] >= 0 ? mirror[j] : j;
            const cd* g2m = gamma + static_cast<std::size_t>(jm) * Nt;
            for (std::size_t t = 0; t < keep; ++t) g2c[t] = std::conj(g2[t]);
            const double wi = omegas[i];
            const double wj = omegas[j];
            for (int k = 0; k < nf; ++k) {
                const double omega = wi + wj + omegas[k];
                const std::size_t off = (static_cast<std::size_t>(idx) * nf + k) * keep;
                if (conv) {
                    ph.build(Nt, omega, dt);
                    for (std::size_t t = 0; t < Nt; ++t) g1_phase[t] = g1[t] * ph.minus[t];
                    padded_spectrum(g1_phase.data(), Nt, Nfft, Fa.data());
                    fc_convolution(g1, g2m, Ghat.data() + static_cast<std::size_t>(jm) * Nfft, ph, Nt,
                                   Nfft, dt, keep, g1_phase.data(), Fa.data(), work.data(), convb.data(),
                                   y.data(), prefix.data(), Fo + off);
                    fc_convolution(g1, g2c.data(), Gchat.data() + static_cast<std::size_t>(j) * Nfft, ph,
                                   Nt, Nfft, dt, keep, g1_phase.data(), Fa.data(), work.data(),
                                   convb.data(), y.data(), prefix.data(), Co + off);
                    r_convolution(g1, g2, ph, keep, dt, y.data(), prefix.data(), prefixP.data(), Ro + off);
                } else {
                    ph.build(keep, omega, dt);
                    f_direct(g1, g2m, ph, keep, dt, false, Fo + off);
                    f_direct(g1, g2, ph, keep, dt, true, Co + off);
                    r_direct(g1, g2, ph, keep, dt, Ro + off);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// MIKX — tcl4_mikx.cpp:46-136 (the code's M1 order), from nf^3 slices at one t.
// slices: s[(a*nf + b)*nf + c]; p2f column-major N x N.
// ---------------------------------------------------------------------------
inline std::size_t f2(const int* p2f, int N, int a, int b) {
    const int v = p2f[a + b * N];
    if (v < 0) throw std::runtime_error("pair_to_freq contains -1 for a required pair");
    return static_cast<std::size_t>(v);
}

void mikx_at(const cd* Fs, const cd* Cs, const cd* Rs, const int* p2f, int N, int nf,
             cd* M, cd* I, cd* K, cd* X) {
    const std::size_t NN = static_cast<std::size_t>(N);
    const std::size_t N2 = NN * NN;
    const std::size_t nf_ = static_cast<std::size_t>(nf);
    auto id3 = [nf_](std::size_t a, std::size_t b, std::size_t c) { return (a * nf_ + b) * nf_ + c; };
    for (std::size_t idx = 0; idx < N2 * N2; ++idx) {
        const std::size_t row_u = idx / N2;
        const std::size_t col_u = idx % N2;
        const int j = static_cast<int>(row_u % NN), k = static_cast<int>(row_u / NN);
        const int p = static_cast<int>(col_u % NN), q = static_cast<int>(col_u / NN);
        const std::size_t f_jk = f2(p2f, N, j, k), f_jq = f2(p2f, N, j, q), f_pj = f2(p2f, N, p, j);
        const std::size_t f_pq = f2(p2f, N, p, q), f_qk = f2(p2f, N, q, k), f_kq = f2(p2f, N, k, q);
        const std::size_t f_qp = f2(p2f, N, q, p), f_qj = f2(p2f, N, q, j);
        const std::size_t e = row_u + col_u * N2;  // (row, col) column-major
        M[e] = Fs[id3(f_jq, f_jk, f_pj)] - Rs[id3(f_jq, f_pq, f_qk)];
        I[e] = Fs[id3(f_jk, f_qp, f_kq)];
        K[e] = Rs[id3(f_jk, f_pq, f_qj)];
        for (int r = 0; r < N; ++r)
            for (int s = 0; s < N; ++s) {
                const std::size_t f_rs = f2(p2f, N, r, s);
                const std::size_t x = static_cast<std::size_t>(j) +
                    NN * (static_cast<std::size_t>(k) + NN * (static_cast<std::size_t>(p) +
                    NN * (static_cast<std::size_t>(q) + NN * (static_cast<std::size_t>(r) +
                    NN * static_cast<std::size_t>(s)))));
                X[x] = Cs[id3(f_jk, f_pq, f_rs)] + Rs[id3(f_jk, f_pq, f_rs)];
            }
    }
}

// Test-only planted mistake: term q of the NAKZWAN sum (1..7 the seven
// off-diagonal terms t1..t7, 8..10 the three j == m terms) enters with the
// wrong sign; 0 (the default) leaves the restatement untouched.  Tests use it
// to show that the closed-form and brute-force checks catch such a mistake.
int g_mutation = 0;
static inline cd mut(int q, cd v) { return g_mutation == q ? -v : v; }

// assemble_liouvillian — tcl4_assemble.cpp:29-178 (NAKZWAN T, then T + T^H).
void assemble_gw(const cd* M, const cd* Ia, const cd* K, const cd* X, const cd* ops, int num_ops,
                 int N, cd* GW) {
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN;
    auto fl2 = [NN](int r, int c) { return static_cast<std::size_t>(r) + static_cast<std::size_t>(c) * NN; };
    auto fl6 = [NN](int j, int k, int p, int q, int r, int s) {
        return static_cast<std::size_t>(j) + NN * (static_cast<std::size_t>(k) + NN * (static_cast<std::size_t>(p) +
               NN * (static_cast<std::size_t>(q) + NN * (static_cast<std::size_t>(r) + NN * static_cast<std::size_t>(s)))));
    };
    auto mat = [N2](const cd* Mx, std::size_t r, std::size_t c) { return Mx[r + c * N2]; };
    for (std::size_t idx = 0; idx < N2 * N2; ++idx) {
        const std::size_t row_u = idx % N2, col_u = idx / N2;
        const int n = static_cast<int>(row_u % NN), i = static_cast<int>(row_u / NN);
        const int m = static_cast<int>(col_u % NN), j = static_cast<int>(col_u / NN);
        cd res(0.0, 0.0);
        for (int iA = 0; iA < num_ops; ++iA) {
            const cd* Ap = ops + static_cast<std::size_t>(iA) * N2;
            auto A = [Ap, NN](int r, int c) { return Ap[static_cast<std::size_t>(r) + static_cast<std::size_t>(c) * NN]; };
            for (int iB = 0; iB < num_ops; ++iB) {
                const cd* Bp = ops + static_cast<std::size_t>(iB) * N2;
                auto B = [Bp, NN](int r, int c) { return Bp[static_cast<std::size_t>(r) + static_cast<std::size_t>(c) * NN]; };
                for (int a = 0; a < N; ++a) {
                    for (int b = 0; b < N; ++b) {
                        const cd M_bi_na = mat(M, fl2(b, i), fl2(n, a));
                        const cd I_an_ib = mat(Ia, fl2(a, n), fl2(i, b));
                        const cd K_ib_na = mat(K, fl2(i, b), fl2(n, a));
                        const cd X_anjbni = X[fl6(a, n, j, b, n, i)];
                        const cd X_iajbni = X[fl6(i, a, j, b, n, i)];
                        const cd X_bajmai = X[fl6(b, a, j, m, a, i)];
                        const cd X_ibjmai = X[fl6(i, b, j, m, a, i)];
                        const cd t1 = B(n, a) * A(a, b) * B(b, i) * A(j, m) * M_bi_na;
                        const cd t2 = A(n, a) * B(a, b) * B(b, i) * A(j, m) * I_an_ib;
                        const cd t3 = B(n, a) * B(a, b) * A(b, i) * A(j, m) * K_ib_na;
                        const cd t4 = A(n, a) * B(a, i) * B(j, b) * A(b, m) * X_anjbni;
                        const cd t5 = B(n, a) * A(a, i) * B(j, b) * A(b, m) * X_iajbni;
                        const cd t6 = A(n, a) * A(a, b) * B(b, i) * B(j, m) * X_bajmai;
                        const cd t7 = A(n, a) * B(a, b) * A(b, i) * B(j, m) * X_ibjmai;
                        res -= (mut(1, t1) - mut(2, t2) + mut(3, t3) + mut(4, t4) - mut(5, t5) - mut(6, t6) +
                                mut(7, t7));
                    }
                }
                if (j == m) {
                    for (int a = 0; a < N; ++a)
                        for (int b = 0; b < N; ++b)
                            for (int c = 0; c < N; ++c) {
                                const cd M_ci_ab = mat(M, fl2(c, i), fl2(a, b));
                                const cd K_ic_ab = mat(K, fl2(i, c), fl2(a, b));
                                const cd I_ba_ic = mat(Ia, fl2(b, a), fl2(i, c));
                                const cd add1 = A(n, a) * B(a, b) * A(b, c) * B(c, i) * M_ci_ab;
                                const cd add2 = A(n, a) * B(a, b) * B(b, c) * A(c, i) * K_ic_ab;
                                const cd add3 = -A(n, a) * A(a, b) * B(b, c) * B(c, i) * I_ba_ic;
                                res += (mut(8, add1) + mut(9, add2) + mut(10, add3));
                            }
                }
            }
        }
        GW[row_u + col_u * N2] = res;
    }
    for (std::size_t col = 0; col < N2; ++col)
        for (std::size_t row = 0; row <= col; ++row) {
            const cd t_rc = GW[row + col * N2];
            const cd t_cr = GW[col + row * N2];
            const cd g = t_rc + std::conj(t_cr);
            GW[row + col * N2] = g;
            if (row != col) GW[col + row * N2] = std::conj(g);
        }
}

// gw_to_liouvillian — tcl4_assemble.cpp:180-205: L4(n,m;i,j) = GW(n,i;m,j).
void gw_to_l4(const cd* GW, int N, cd* L) {
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN;
    for (std::size_t n = 0; n < NN; ++n)
        for (std::size_t m = 0; m < NN; ++m)
            for (std::size_t i = 0; i < NN; ++i)
                for (std::size_t j = 0; j < NN; ++j)
                    L[(n + m * NN) + (i + j * NN) * N2] = GW[(n + i * NN) + (m + j * NN) * N2];
}

// One L4(t) from nf^3 slices at that t.
void l4_from_slices(const cd* Fs, const cd* Cs, const cd* Rs, const int* p2f, int N, int nf,
                    const cd* ops, int num_ops, cd* L4) {
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN, N4 = N2 * N2, N6 = N4 * N2;
    std::vector<cd> M(N4), I(N4), K(N4), X(N6), GW(N4);
    mikx_at(Fs, Cs, Rs, p2f, N, nf, M.data(), I.data(), K.data(), X.data());
    assemble_gw(M.data(), I.data(), K.data(), X.data(), ops, num_ops, N, GW.data());
    gw_to_l4(GW.data(), N, L4);
}

// Gather the nf^3 slice at time t from series-major kernels.
void slice_at(const cd* S, std::size_t nf3, std::size_t keep, std::size_t t, cd* out) {
    for (std::size_t e = 0; e < nf3; ++e) out[e] = S[e * keep + t];
}

// ---------------------------------------------------------------------------
// TCL2 superoperator — generator.cpp:42-177 with the driver's kernels
// (TCL4_spin_boson_example.cpp:293-322): Gamma(a,a) = gamma(t,b), off-diag 0.
// ---------------------------------------------------------------------------
double sqnorm(const std::vector<cd>& A) {
    double s = 0.0;
    for (const auto& v : A) s += std::norm(v);
    return s;
}

// TCL2 generator (build_tcl2_components, generator.cpp:81-177) from channel
// matrices: gamma_ab[(b*num_ops + beta)*num_ops + alpha] = Gamma_b(alpha, beta) at t.
// tcl2_at (below) is the drivers' diagonal case Gamma_b(alpha, beta) =
// delta_ab gamma_b (TCL4_spin_boson_example.cpp:303-316).
void tcl2_general_at(const double* eps, const cd* ops_eig, int num_ops, int N, const int* p2f, int nf,
                     const cd* gamma_ab, cd* Lout) {
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN;
    std::vector<cd> L(N2 * N2, cd(0.0, 0.0)), Hls(N2, cd(0.0, 0.0));
    for (int b = 0; b < nf; ++b) {
        std::vector<std::vector<cd>> parts(static_cast<std::size_t>(num_ops));
        std::vector<char> has(static_cast<std::size_t>(num_ops), 0);
        for (int al = 0; al < num_ops; ++al) {
            std::vector<cd> P(N2, cd(0.0, 0.0));
            const cd* A = ops_eig + static_cast<std::size_t>(al) * N2;
            for (int mm = 0; mm < N; ++mm)
                for (int n2 = 0; n2 < N; ++n2)
                    if (p2f[mm + n2 * N] == b) P[mm + n2 * NN] = A[mm + n2 * NN];
            has[al] = sqnorm(P) != 0.0;
            parts[al] = std::move(P);
        }
        for (int al = 0; al < num_ops; ++al) {
            if (!has[al]) continue;
            for (int be = 0; be < num_ops; ++be) {
                if (!has[be]) continue;
                const cd G = gamma_ab[(static_cast<std::size_t>(b) * num_ops + be) * num_ops + al];
                const double gam = 2.0 * G.real();
                const double S = G.imag();
                const bool need_jump = std::abs(gam) > 0.0;
                const bool need_ls = std::abs(S) > 0.0;
                if (!need_jump && !need_ls) continue;
                const auto& Aa = parts[al];
                const auto& Ab = parts[be];
                // AdagA = Aa^dagger * Ab
                std::vector<cd> AdA(N2, cd(0.0, 0.0));
                for (std::size_t r = 0; r < NN; ++r)
                    for (std::size_t c = 0; c < NN; ++c) {
                        cd s(0.0, 0.0);
                        for (std::size_t q = 0; q < NN; ++q) s += std::conj(Aa[q + r * NN]) * Ab[q + c * NN];
                        AdA[r + c * NN] = s;
                    }
                if (sqnorm(AdA) == 0.0) continue;
                if (need_ls)
                    for (std::size_t e = 0; e < N2; ++e) Hls[e] += S * AdA[e];
                if (!need_jump) continue;
                // L += gam*(kron(conj(Aa), Ab) - 0.5 kron(I, AdA) - 0.5 kron(AdA^T, I))
                for (std::size_t c1 = 0; c1 < NN; ++c1)
                    for (std::size_t c2 = 0; c2 < NN; ++c2)
                        for (std::size_t r1 = 0; r1 < NN; ++r1)
                            for (std::size_t r2 = 0; r2 < NN; ++r2) {
                                const std::size_t row = r1 * NN + r2, col = c1 * NN + c2;
                                const cd jump = std::conj(Aa[r1 + c1 * NN]) * Ab[r2 + c2 * NN];
                                const cd left = (r1 == c1) ? AdA[r2 + c2 * NN] : cd(0.0, 0.0);
                                const cd right = (r2 == c2) ? AdA[c1 + r1 * NN] : cd(0.0, 0.0);
                                L[row + col * N2] += gam * (jump - 0.5 * left - 0.5 * right);
                            }
            }
        }
    }
    // H_eff = diag(eps) + hermitize(H_ls); L_unitary = -i (I (x) H - H^T (x) I)
    std::vector<cd> H(N2);
    for (std::size_t r = 0; r < NN; ++r)
        for (std::size_t c = 0; c < NN; ++c)
            H[r + c * NN] = 0.5 * (Hls[r + c * NN] + std::conj(Hls[c + r * NN])) + (r == c ? cd(eps[r], 0.0) : cd(0.0, 0.0));
    const cd mi(0.0, -1.0);
    for (std::size_t c1 = 0; c1 < NN; ++c1)
        for (std::size_t c2 = 0; c2 < NN; ++c2)
            for (std::size_t r1 = 0; r1 < NN; ++r1)
                for (std::size_t r2 = 0; r2 < NN; ++r2) {
                    const std::size_t row = r1 * NN + r2, col = c1 * NN + c2;
                    const cd left = (r1 == c1) ? H[r2 + c2 * NN] : cd(0.0, 0.0);
                    const cd right = (r2 == c2) ? H[c1 + r1 * NN] : cd(0.0, 0.0);
                    Lout[row + col * N2] = mi * (left - right) + L[row + col * N2];
                }
}

void tcl2_at(const double* eps, const cd* ops_eig, int num_ops, int N, const int* p2f, int nf,
             const cd* gamma_row /* nf values at t */, cd* Lout) {
    const std::size_t P = static_cast<std::size_t>(num_ops) * num_ops;
    std::vector<cd> gab(static_cast<std::size_t>(nf) * P, cd(0.0, 0.0));
    for (int b = 0; b < nf; ++b)
        for (int a = 0; a < num_ops; ++a) gab[b * P + static_cast<std::size_t>(a) * num_ops + a] = gamma_row[b];
    tcl2_general_at(eps, ops_eig, num_ops, N, p2f, nf, gab.data(), Lout);
}

// ---------------------------------------------------------------------------
// RK4 — rk4_dense.cpp:84-160 (matvec_serial, rk4_update_serial).
// ---------------------------------------------------------------------------
void matvec(const cd* L, const cd* x, std::size_t D, cd* y) {
    for (std::size_t i = 0; i < D; ++i) {
        cd sum(0.0, 0.0);
        for (std::size_t j = 0; j < D; ++j) sum += L[i + j * D] * x[j];
        y[i] = sum;
    }
}

void rk4_update(const cd* L0, const cd* Lh, const cd* L1, cd* r, std::size_t D, double dt) {
    std::vector<cd> k1(D), k2(D), k3(D), k4(D), tmp(D);
    matvec(L0, r, D, k1.data());
    for (std::size_t i = 0; i < D; ++i) tmp[i] = r[i] + (0.5 * dt) * k1[i];
    matvec(Lh, tmp.data(), D, k2.data());
    for (std::size_t i = 0; i < D; ++i) tmp[i] = r[i] + (0.5 * dt) * k2[i];
    matvec(Lh, tmp.data(), D, k3.data());
    for (std::size_t i = 0; i < D; ++i) tmp[i] = r[i] + dt * k3[i];
    matvec(L1, tmp.data(), D, k4.data());
    const double s = dt / 6.0;
    for (std::size_t i = 0; i < D; ++i) r[i] += s * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

// ---------------------------------------------------------------------------
// Hermitian eigensolver (cyclic complex Jacobi) standing in for Eigen's
// SelfAdjointEigenSolver (system.hpp:44).  Eigenvalues ascending; each
// eigenvector's largest-magnitude entry is made real positive.
// ---------------------------------------------------------------------------
void jacobi_eigh(const cd* Hin, int n, double* w, cd* V) {
    const std::size_t N = static_cast<std::size_t>(n);
    std::vector<cd> A(Hin, Hin + N * N);
    for (std::size_t r = 0; r < N; ++r)
        for (std::size_t c = 0; c < N; ++c) V[r + c * N] = (r == c) ? cd(1.0, 0.0) : cd(0.0, 0.0);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (std::size_t c = 0; c < N; ++c)
            for (std::size_t r = 0; r < N; ++r) {
                tot += std::norm(A[r + c * N]);
                if (r != c) off += std::norm(A[r + c * N]);
            }
        if (off <= 1e-30 * std::max(tot, 1e-300)) break;
        for (std::size_t p = 0; p < N; ++p)
            for (std::size_t q = p + 1; q < N; ++q) {
                const cd apq = A[p + q * N];
                const double mag = std::abs(apq);
                if (mag < 1e-300) continue;
                const double app = A[p + p * N].real(), aqq = A[q + q * N].real();
                const cd ph = apq / mag;  // e^{i phi}
                const double theta = 0.5 * std::atan2(2.0 * mag, aqq - app);
                const double c = std::cos(theta), s = std::sin(theta);
                // Rotation G with columns: p -> c e_p - s conj(ph)... applied as A <- G^H A G
                // G = [[c, s*ph],[-s*conj(ph), c]] acting on (p,q) — chosen to zero A(p,q).
                const cd gpp(c, 0.0), gpq = s * ph, gqp = -s * std::conj(ph), gqq(c, 0.0);
                // A <- A G (columns p,q)
                for (std::size_t r = 0; r < N; ++r) {
                    const cd arp = A[r + p * N], arq = A[r + q * N];
                    A[r + p * N] = arp * gpp + arq * gqp;
                    A[r + q * N] = arp * gpq + arq * gqq;
                }
                // A <- G^H A (rows p,q)
                for (std::size_t cc = 0; cc < N; ++cc) {
                    const cd apc = A[p + cc * N], aqc = A[q + cc * N];
                    A[p + cc * N] = std::conj(gpp) * apc + std::conj(gqp) * aqc;
                    A[q + cc * N] = std::conj(gpq) * apc + std::conj(gqq) * aqc;
                }
                A[p + q * N] = cd(0.0, 0.0);
                A[q + p * N] = cd(0.0, 0.0);
                A[p + p * N] = cd(A[p + p * N].real(), 0.0);
                A[q + q * N] = cd(A[q + q * N].real(), 0.0);
                for (std::size_t r = 0; r < N; ++r) {
                    const cd vrp = V[r + p * N], vrq = V[r + q * N];
                    V[r + p * N] = vrp * gpp + vrq * gqp;
                    V[r + q * N] = vrp * gpq + vrq * gqq;
                }
            }
    }
    std::vector<std::size_t> order(N);
    for (std::size_t i = 0; i < N; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](std::size_t a, std::size_t b) { return A[a + a * N].real() < A[b + b * N].real(); });
    std::vector<cd> Vs(N * N);
    for (std::size_t c = 0; c < N; ++c) {
        const std::size_t src = order[c];
        w[c] = A[src + src * N].real();
        std::size_t imax = 0;
        double vmax = -1.0;
        for (std::size_t r = 0; r < N; ++r)
            if (std::abs(V[r + src * N]) > vmax + 1e-14) { vmax = std::abs(V[r + src * N]); imax = r; }
        const cd v0 = V[imax + src * N];
        const cd phase = (std::abs(v0) > 0) ? std::conj(v0) / std::abs(v0) : cd(1.0, 0.0);
        double nrm = 0.0;
        for (std::size_t r = 0; r < N; ++r) nrm += std::norm(V[r + src * N]);
        nrm = std::sqrt(nrm);
        for (std::size_t r = 0; r < N; ++r) Vs[r + c * N] = V[r + src * N] * phase / nrm;
    }
    std::copy(Vs.begin(), Vs.end(), V);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// 0: std::exp phases (glibc on this platform); 1: correctly rounded phases.
void orc_set_phase_mode(int mode) { g_phase_mode = mode; }
void orc_set_mutation(int term) { g_mutation = term; }
int orc_get_phase_mode() { return g_phase_mode; }

// Phases exp(i*omega*(k*dt)), k < n, in the current mode (interleaved re, im).
void orc_phases(double omega, double dt, size_t n, double* out) {
    for (std::size_t k = 0; k < n; ++k) {
        const cd p = phase(omega * (static_cast<double>(k) * dt));
        out[2 * k] = p.real();
        out[2 * k + 1] = p.imag();
    }
}

// Correctly rounded sin/cos of each x (for pinning against quad precision).
void orc_cr_sincos(const double* x, size_t n, double* s, double* c) {
    for (std::size_t k = 0; k < n; ++k) cr_sincos(x[k], s + k, c + k);
}

// Correctly rounded exp of each x (quad precision rounded to double).
void orc_cr_exp(const double* x, size_t n, double* out) {
    for (std::size_t k = 0; k < n; ++k) out[k] = static_cast<double>(expq(static_cast<__float128>(x[k])));
}

}  // extern "C"

namespace {
template <class Fn>
int guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown error";
        return 1;
    }
}
}  // namespace

extern "C" {


size_t orc_next_pow2(size_t n) { return next_pow2(n); }
size_t orc_fcr_nfft(size_t Nt, size_t pad) { return fcr_nfft(Nt, pad); }

int orc_fft(double* a, size_t n, int forward) {
    return guard([&] {fft_ref(reinterpret_cast<cd*>(a), n, forward != 0);});
}

int orc_bcf_fft(size_t N, double dt, int kind, double p0, double p1, double beta, double* C_out) {
    return guard([&] {bcf_fft(N, dt, kind, p0, p1, beta, reinterpret_cast<cd*>(C_out));});
}

int orc_gamma_trapz(const double* C, size_t N, double dt, const double* omegas, int M, double* G) {
    return guard([&] {gamma_trapz(reinterpret_cast<const cd*>(C), N, dt, omegas, M, reinterpret_cast<cd*>(G));});
}

int orc_gamma_rect(const double* C, size_t N, double dt, const double* omegas, int M, double* G) {
    return guard([&] {gamma_rect(reinterpret_cast<const cd*>(C), N, dt, omegas, M, reinterpret_cast<cd*>(G));});
}

// Scalar F/C/R series of one triple (compute_F/C/R_series, tcl4_kernels.cpp:283-320).
// which: 0 F, 1 C, 2 R; method: 0 Convolution, 1 Direct.
int orc_fcr_series(const double* g1, const double* g2, size_t Nt, double omega, double dt,
                   int which, int method, size_t pad, double* out) {
    return guard([&] {
        const cd* a = reinterpret_cast<const cd*>(g1);
        const cd* b = reinterpret_cast<const cd*>(g2);
        cd* o = reinterpret_cast<cd*>(out);
        Phases ph;
        ph.build(Nt, omega, dt);
        if (method == 1) {
            if (which == 2) r_direct(a, b, ph, Nt, dt, o);
            else f_direct(a, b, ph, Nt, dt, which == 1, o);
        } else if (which == 2) {
            std::vector<cd> y(Nt), p(Nt), pp(Nt);
            r_convolution(a, b, ph, Nt, dt, y.data(), p.data(), pp.data(), o);
        } else {
            const std::size_t Nfft = fcr_nfft(Nt, pad);
            std::vector<cd> g1p(Nt), op(Nt), Fa(Nfft), Gb(Nfft), work(Nfft), conv(Nt), y(Nt), pre(Nt);
            for (std::size_t t = 0; t < Nt; ++t) {
                g1p[t] = a[t] * ph.minus[t];
                op[t] = (which == 1) ? std::conj(b[t]) : b[t];
            }
            padded_spectrum(g1p.data(), Nt, Nfft, Fa.data());
            padded_spectrum(op.data(), Nt, Nfft, Gb.data());
            fc_convolution(a, op.data(), Gb.data(), ph, Nt, Nfft, dt, Nt, g1p.data(), Fa.data(),
                           work.data(), conv.data(), y.data(), pre.data(), o);
        }
    });
}

// causal_conv_fft (tcl4_kernels.cpp:31-68) on its own: out = dt*ifft(fft(f).*fft(g))[0..N).
int orc_causal_conv_fft(const double* f, const double* g, size_t N, double dt, size_t pad, double* out) {
    return guard([&] {
        const std::size_t Nfft = fcr_nfft(N, pad);
        std::vector<cd> Fa(Nfft), Gb(Nfft), work(Nfft);
        padded_spectrum(reinterpret_cast<const cd*>(f), N, Nfft, Fa.data());
        padded_spectrum(reinterpret_cast<const cd*>(g), N, Nfft, Gb.data());
        conv_from_spectra(Fa.data(), Gb.data(), Nfft, dt, N, work.data(), reinterpret_cast<cd*>(out));
    });
}

int orc_mirror(const double* omegas, int nf, double fidx_tol, int* mirror) {
    return guard([&] {mirror_indices(omegas, nf, fidx_tol, mirror);});
}

int orc_triple_kernels(const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                       double dt, int method, size_t pad, size_t keep, double* F, double* C, double* R,
                       int nthreads) {
    return guard([&] {triple_kernels(reinterpret_cast<const cd*>(gamma), Nt, nf, omegas, mirror, dt, method, pad, keep,
                           reinterpret_cast<cd*>(F), reinterpret_cast<cd*>(C), reinterpret_cast<cd*>(R),
                           nthreads > 0 ? nthreads : 1);});
}

int orc_mikx(const double* Fs, const double* Cs, const double* Rs, const int* p2f, int N, int nf,
             double* M, double* I, double* K, double* X) {
    return guard([&] {mikx_at(reinterpret_cast<const cd*>(Fs), reinterpret_cast<const cd*>(Cs),
                    reinterpret_cast<const cd*>(Rs), p2f, N, nf, reinterpret_cast<cd*>(M),
                    reinterpret_cast<cd*>(I), reinterpret_cast<cd*>(K), reinterpret_cast<cd*>(X));});
}

int orc_assemble_gw(const double* M, const double* I, const double* K, const double* X, const double* ops,
                    int num_ops, int N, double* GW) {
    return guard([&] {
        if (num_ops <= 0) throw std::invalid_argument("assemble_liouvillian: coupling_ops must be non-empty");
        assemble_gw(reinterpret_cast<const cd*>(M), reinterpret_cast<const cd*>(I),
                    reinterpret_cast<const cd*>(K), reinterpret_cast<const cd*>(X),
                    reinterpret_cast<const cd*>(ops), num_ops, N, reinterpret_cast<cd*>(GW));
    });
}

int orc_gw_to_l4(const double* GW, int N, double* L) {
    return guard([&] {gw_to_l4(reinterpret_cast<const cd*>(GW), N, reinterpret_cast<cd*>(L));});
}

int orc_l4_from_slices(const double* Fs, const double* Cs, const double* Rs, const int* p2f, int N, int nf,
                       const double* ops, int num_ops, double* L4) {
    return guard([&] {l4_from_slices(reinterpret_cast<const cd*>(Fs), reinterpret_cast<const cd*>(Cs),
                           reinterpret_cast<const cd*>(Rs), p2f, N, nf, reinterpret_cast<const cd*>(ops),
                           num_ops, reinterpret_cast<cd*>(L4));});
}

// End-to-end L4 at a list of time indices (build_correction_series restricted
// to tidx; tcl4_e2e_cuda_compare.cpp:183-202 CPU leg).  Kernels are kept only
// for t <= max(tidx).  L4_out: ntau blocks of N^4, column-major.
int orc_l4_series(const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                  const int* p2f, int N, const double* ops, int num_ops, double dt, int method, size_t pad,
                  const size_t* tidx, size_t ntau, double* L4_out, int nthreads) {
    return guard([&] {
        if (ntau == 0) return;
        std::size_t tmax = 0;
        for (std::size_t q = 0; q < ntau; ++q) {
            if (tidx[q] >= Nt) throw std::out_of_range("time_index out of range");
            tmax = std::max(tmax, tidx[q]);
        }
        const std::size_t keep = tmax + 1;
        const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
        std::vector<cd> F(nf3 * keep), C(nf3 * keep), R(nf3 * keep);
        const int nt = nthreads > 0 ? nthreads : 1;
        triple_kernels(reinterpret_cast<const cd*>(gamma), Nt, nf, omegas, mirror, dt, method, pad, keep,
                       F.data(), C.data(), R.data(), nt);
        const std::size_t N4 = static_cast<std::size_t>(N) * N * N * N;
        std::string err;
        #pragma omp parallel num_threads(nt)
        {
            std::vector<cd> Fs(nf3), Cs(nf3), Rs(nf3);
            #pragma omp for schedule(dynamic)
            for (std::ptrdiff_t q = 0; q < static_cast<std::ptrdiff_t>(ntau); ++q) {
                try {
                    slice_at(F.data(), nf3, keep, tidx[q], Fs.data());
                    slice_at(C.data(), nf3, keep, tidx[q], Cs.data());
                    slice_at(R.data(), nf3, keep, tidx[q], Rs.data());
                    l4_from_slices(Fs.data(), Cs.data(), Rs.data(), p2f, N, nf, reinterpret_cast<const cd*>(ops),
                                   num_ops, reinterpret_cast<cd*>(L4_out) + static_cast<std::size_t>(q) * N4);
                } catch (const std::exception& e) {
                    #pragma omp critical
                    err = e.what();
                }
            }
        }
        if (!err.empty()) throw std::runtime_error(err);
    });
}

int orc_tcl2_at(const double* eps, const double* ops_eig, int num_ops, int N, const int* p2f, int nf,
                const double* gamma_row, double* L) {
    return guard([&] {tcl2_at(eps, reinterpret_cast<const cd*>(ops_eig), num_ops, N, p2f, nf,
                    reinterpret_cast<const cd*>(gamma_row), reinterpret_cast<cd*>(L));});
}

int orc_tcl2_general_at(const double* eps, const double* ops_eig, int num_ops, int N, const int* p2f, int nf,
                        const double* gamma_ab, double* L) {
    return guard([&] {tcl2_general_at(eps, reinterpret_cast<const cd*>(ops_eig), num_ops, N, p2f, nf,
                    reinterpret_cast<const cd*>(gamma_ab), reinterpret_cast<cd*>(L));});
}

int orc_rk4_update(const double* L0, const double* Lh, const double* L1, double* r, size_t D, double dt) {
    return guard([&] {
        if (!(dt > 0.0)) throw std::invalid_argument("rk4_update_serial: dt must be > 0");
        rk4_update(reinterpret_cast<const cd*>(L0), reinterpret_cast<const cd*>(Lh), reinterpret_cast<const cd*>(L1),
                   reinterpret_cast<cd*>(r), D, dt);
    });
}

// propagate_rk4_dense_serial(L_series[, L_half_series], r, t0, dt, on_sample, every)
// rk4_dense.cpp:246-320.  samples_out receives r at steps 0, every, 2*every...
// plus the final state (the reference's on_sample calls).  Returns the count.
int orc_propagate(const double* L_series, const double* L_half, size_t steps, size_t D, double* r,
                  double dt, size_t every, double* samples_out, size_t* nsamples) {
    return guard([&] {
        if (!(dt > 0.0)) throw std::invalid_argument("propagate_rk4_dense_serial: dt must be > 0");
        if (steps < 1) throw std::invalid_argument("L_series must have at least 2 matrices (endpoints)");
        if (every == 0) throw std::invalid_argument("sample_every must be > 0");
        const cd* Ls = reinterpret_cast<const cd*>(L_series);
        const cd* Lh = reinterpret_cast<const cd*>(L_half);
        cd* rv = reinterpret_cast<cd*>(r);
        cd* so = reinterpret_cast<cd*>(samples_out);
        const std::size_t D2 = D * D;
        std::vector<cd> half(D2);
        std::size_t ns = 0;
        for (std::size_t i = 0; i < steps; ++i) {
            if (so && i % every == 0) { std::copy(rv, rv + D, so + ns * D); ++ns; }
            const cd* L0 = Ls + i * D2;
            const cd* L1 = Ls + (i + 1) * D2;
            const cd* H;
            if (Lh) {
                H = Lh + i * D2;
            } else {
                for (std::size_t e = 0; e < D2; ++e) half[e] = 0.5 * (L0[e] + L1[e]);
                H = half.data();
            }
            rk4_update(L0, H, L1, rv, D, dt);
        }
        if (so) { std::copy(rv, rv + D, so + ns * D); ++ns; }
        if (nsamples) *nsamples = ns;
    });
}

// Driver composition of TCL4_spin_boson_example.cpp:284-384: L(t) = TCL2(t)
// [+ L4(t) when order == 4]; order 0 uses L0 = -i[diag(eps), .] only.
// Midpoint generator 0.5*(L_n + L_{n+1}); r_inout is vec(rho) in the eigenbasis.
// samples_out: (steps+1) x N^2 states (r before each step and the final one).
int orc_tcl4_propagate(const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                       const int* p2f, int N, const double* eps, const double* ops, int num_ops, double dt,
                       int method, size_t pad, size_t steps, int order, double* r_inout, double* samples_out,
                       int nthreads) {
    return guard([&] {
        if (!(dt > 0.0)) throw std::invalid_argument("dt must be > 0");
        if (!(order == 0 || order == 2 || order == 4)) throw std::invalid_argument("order must be 0, 2, or 4");
        if (steps + 1 > Nt) throw std::out_of_range("steps exceed gamma_series rows");
        const std::size_t NN = static_cast<std::size_t>(N), D = NN * NN, D2 = D * D;
        const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
        const std::size_t keep = steps + 1;
        const int nt = nthreads > 0 ? nthreads : 1;
        std::vector<cd> F, C, R;
        if (order == 4) {
            F.resize(nf3 * keep); C.resize(nf3 * keep); R.resize(nf3 * keep);
            triple_kernels(reinterpret_cast<const cd*>(gamma), Nt, nf, omegas, mirror, dt, method, pad, keep,
                           F.data(), C.data(), R.data(), nt);
        }
        const cd* G = reinterpret_cast<const cd*>(gamma);
        const cd* opsc = reinterpret_cast<const cd*>(ops);
        std::vector<cd> Lall(keep * D2);
        std::string err;
        #pragma omp parallel num_threads(nt)
        {
            std::vector<cd> Fs(nf3), Cs(nf3), Rs(nf3), L4(D2), grow(static_cast<std::size_t>(nf));
            #pragma omp for schedule(dynamic)
            for (std::ptrdiff_t tt = 0; tt < static_cast<std::ptrdiff_t>(keep); ++tt) {
                try {
                    const std::size_t t = static_cast<std::size_t>(tt);
                    cd* L = Lall.data() + t * D2;
                    if (order == 0) {
                        for (std::size_t e = 0; e < D2; ++e) L[e] = cd(0.0, 0.0);
                        const cd mi(0.0, -1.0);
                        for (std::size_t c1 = 0; c1 < NN; ++c1)
                            for (std::size_t c2 = 0; c2 < NN; ++c2) {
                                const std::size_t row = c1 * NN + c2;
                                const cd left(eps[c2], 0.0), right(eps[c1], 0.0);
                                L[row + row * D] = mi * (left - right);
                            }
                        continue;
                    }
                    for (int b = 0; b < nf; ++b) grow[static_cast<std::size_t>(b)] = G[t + static_cast<std::size_t>(b) * Nt];
                    tcl2_at(eps, opsc, num_ops, N, p2f, nf, grow.data(), L);
                    if (order == 4) {
                        slice_at(F.data(), nf3, keep, t, Fs.data());
                        slice_at(C.data(), nf3, keep, t, Cs.data());
                        slice_at(R.data(), nf3, keep, t, Rs.data());
                        l4_from_slices(Fs.data(), Cs.data(), Rs.data(), p2f, N, nf, opsc, num_ops, L4.data());
                        for (std::size_t e = 0; e < D2; ++e) L[e] += L4[e];
                    }
                } catch (const std::exception& e) {
                    #pragma omp critical
                    err = e.what();
                }
            }
        }
        if (!err.empty()) throw std::runtime_error(err);
        cd* r = reinterpret_cast<cd*>(r_inout);
        cd* so = reinterpret_cast<cd*>(samples_out);
        std::vector<cd> half(D2);
        if (so) std::copy(r, r + D, so);
        for (std::size_t s = 0; s < steps; ++s) {
            const cd* L0 = Lall.data() + s * D2;
            const cd* L1 = Lall.data() + (s + 1) * D2;
            for (std::size_t e = 0; e < D2; ++e) half[e] = 0.5 * (L0[e] + L1[e]);
            rk4_update(L0, half.data(), L1, r, D, dt);
            if (so) std::copy(r, r + D, so + (s + 1) * D);
        }
    });
}

// sys::System::build (system.hpp:212-225) + build_map (tcl4.cpp:40-101).
// Outputs sized for nf <= N^2.  U column-major N x N; A_eig num_ops x N^2.
int orc_system_build(const double* H, int N, const double* A_lab, int num_ops, double freq_tol, double* eps,
                     double* U, double* A_eig, int* nf_out, double* omegas, int* mirror, int* p2f,
                     int* zero_index) {
    return guard([&] {
        if (N <= 0) throw std::invalid_argument("Eigensystem: Hamiltonian dimension must be > 0");
        const std::size_t NN = static_cast<std::size_t>(N);
        cd* Uc = reinterpret_cast<cd*>(U);
        jacobi_eigh(reinterpret_cast<const cd*>(H), N, eps, Uc);
        const double tol = std::max(freq_tol, std::numeric_limits<double>::epsilon());
        struct Bucket { double omega; std::vector<std::pair<int, int>> pairs; };
        std::vector<Bucket> bk;
        for (int m = 0; m < N; ++m)
            for (int n = 0; n < N; ++n) {
                const double w = eps[m] - eps[n];
                std::size_t idx = bk.size();
                for (std::size_t q = 0; q < bk.size(); ++q)
                    if (std::abs(bk[q].omega - w) <= tol) { idx = q; break; }
                if (idx == bk.size()) bk.push_back(Bucket{w, {}});
                bk[idx].pairs.emplace_back(m, n);
            }
        std::sort(bk.begin(), bk.end(), [](const Bucket& a, const Bucket& b) { return a.omega < b.omega; });
        const int nf = static_cast<int>(bk.size());
        *nf_out = nf;
        for (std::size_t e = 0; e < NN * NN; ++e) p2f[e] = -1;
        for (int b = 0; b < nf; ++b) {
            omegas[b] = bk[static_cast<std::size_t>(b)].omega;
            for (const auto& pr : bk[static_cast<std::size_t>(b)].pairs) p2f[pr.first + pr.second * N] = b;
        }
        mirror_indices(omegas, nf, tol, mirror);
        const double mtol = std::max(1e-12, tol);
        int zi = -1;
        for (int b = 0; b < nf; ++b)
            if (std::abs(omegas[b]) <= mtol) { zi = b; break; }
        if (zi < 0) {
            double best = std::numeric_limits<double>::infinity();
            for (int b = 0; b < nf; ++b)
                if (std::abs(omegas[b]) < best) { best = std::abs(omegas[b]); zi = b; }
        }
        *zero_index = zi;
        // A_eig = U^dagger A U
        const cd* Al = reinterpret_cast<const cd*>(A_lab);
        cd* Ae = reinterpret_cast<cd*>(A_eig);
        std::vector<cd> tmp(NN * NN);
        for (int o = 0; o < num_ops; ++o) {
            const cd* A = Al + static_cast<std::size_t>(o) * NN * NN;
            cd* E = Ae + static_cast<std::size_t>(o) * NN * NN;
            for (std::size_t r = 0; r < NN; ++r)
                for (std::size_t c = 0; c < NN; ++c) {
                    cd s(0.0, 0.0);
                    for (std::size_t q = 0; q < NN; ++q) s += std::conj(Uc[q + r * NN]) * A[q + c * NN];
                    tmp[r + c * NN] = s;
                }
            for (std::size_t r = 0; r < NN; ++r)
                for (std::size_t c = 0; c < NN; ++c) {
                    cd s(0.0, 0.0);
                    for (std::size_t q = 0; q < NN; ++q) s += tmp[r + q * NN] * Uc[q + c * NN];
                    E[r + c * NN] = s;
                }
        }
    });
}

}  // extern "C"
