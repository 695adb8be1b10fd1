// host_prep.cpp — host-side preparation of the TCL4 hot-path inputs.
//
//  * system build: Hermitian eigensystem, Bohr-frequency buckets, pair map,
//    mirror and zero bucket (system.hpp:37-157, 201-226; tcl4.cpp:40-101);
//  * C(t) from a spectral density (bcf_fft_fun, correlation_fft.hpp:154-214);
//  * the trapezoid Gamma table (gamma.hpp:166-213);
//  * the recurrence twiddle tables consumed by the GPU FFT
//    (correlation_fft.hpp:52-66).
//
// Rounding-sensitive steps keep the reference's operation order; the file
// is compiled with -ffp-contract=off so no product is fused into an add.

#include "engine_internal.hpp"
#include "cr_exp.hpp"

#include <quadmath.h>

#include "cr_sincos.hpp"
#include "system_prep.hpp"

#include <algorithm>
#include <cstdint>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cmath>
#include <complex>
#include <functional>
#include <limits>
#include <stdexcept>
#include <cstring>
#include <vector>

namespace taco_engine {

using cd = std::complex<double>;
static constexpr double kPi = 3.141592653589793238462643383279502884;

std::size_t next_pow2(std::size_t n) {
    if (n == 0) return 1;
    --n;
    n |= n >> 1;  n |= n >> 2;  n |= n >> 4;
    n |= n >> 8;  n |= n >> 16; n |= n >> 32;
    return ++n;
}

std::size_t fcr_nfft(std::size_t Nt, std::size_t pad) {
    const std::size_t L = 2 * Nt - 1;
    std::size_t target = L;
    if (pad > 0) target = std::max(L, pad * Nt);
    std::size_t n = next_pow2(target);
    return n < 2 ? 2 : n;
}

// Stage twiddles of the radix-2 FFT of length 2^L: for stage s (block length
// 2^s) entry k < 2^(s-1) is the k-th power of wlen = exp(-2 pi i / 2^s)
// obtained by repeated complex multiplication from 1, exactly the sequence
// the reference's inner loop produces (w *= wlen restarts at 1 per block).
// Table offset of stage s is 2^(s-1) - 1; total 2^L - 1 entries.
void make_twiddles(int L, std::vector<double>& out) {
    const std::size_t total = (std::size_t{1} << L) - 1;
    out.assign(2 * total, 0.0);
    for (int s = 1; s <= L; ++s) {
        const std::size_t len = std::size_t{1} << s;
        const double ang = (-1.0) * 2.0 * kPi / static_cast<double>(len);
        const double lr = std::cos(ang), li = std::sin(ang);
        double wr = 1.0, wi = 0.0;
        const std::size_t off = (len >> 1) - 1;
        for (std::size_t k = 0; k < (len >> 1); ++k) {
            out[2 * (off + k)] = wr;
            out[2 * (off + k) + 1] = wi;
            const double nr = wr * lr - wi * li;
            const double ni = wr * li + wi * lr;
            wr = nr;
            wi = ni;
        }
    }
}

// Host radix-2 FFT in the reference's form (bit reversal, recurrence twiddles).
static void fft_host(std::vector<cd>& a, bool forward) {
    const std::size_t n = a.size();
    if (n == 0) return;
    std::size_t j = 0;
    for (std::size_t i = 1; i < n; ++i) {
        std::size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (std::size_t len = 2; len <= n; len <<= 1) {
        const double ang = (forward ? -1.0 : +1.0) * 2.0 * kPi / static_cast<double>(len);
        const double lr = std::cos(ang), li = std::sin(ang);
        const std::size_t half = len >> 1;
        for (std::size_t i = 0; i < n; i += len) {
            double wr = 1.0, wi = 0.0;
            for (std::size_t k = 0; k < half; ++k) {
                const cd u = a[i + k];
                const cd x = a[i + k + half];
                const cd v(wr * x.real() - wi * x.imag(), wr * x.imag() + wi * x.real());
                a[i + k] = cd(u.real() + v.real(), u.imag() + v.imag());
                a[i + k + half] = cd(u.real() - v.real(), u.imag() - v.imag());
                const double nr = wr * lr - wi * li;
                const double ni = wr * li + wi * lr;
                wr = nr;
                wi = ni;
            }
        }
    }
    if (!forward) {
        const double inv_n = 1.0 / static_cast<double>(n);
        for (auto& x : a) x = cd(x.real() * inv_n, x.imag() * inv_n);
    }
}

void make_twiddles_exact(int L, std::vector<double>& out) {
    const std::size_t total = (std::size_t{1} << L) - 1;
    out.assign(2 * total, 0.0);
    const __float128 two_pi = 2 * acosq(static_cast<__float128>(-1));
    for (int s = 1; s <= L; ++s) {
        const std::size_t half = std::size_t{1} << (s - 1), off = half - 1;
        for (std::size_t k = 0; k < half; ++k) {
            const __float128 ang = -two_pi * static_cast<__float128>(k) / static_cast<__float128>(2 * half);
            out[2 * (off + k)] = static_cast<double>(cosq(ang));
            out[2 * (off + k) + 1] = static_cast<double>(sinq(ang));
        }
    }
}

static int g_fft_mode = TACO_FFT_PARITY;

int fft_mode() { return g_fft_mode; }

void set_fft_mode(int mode) {
    if (mode != TACO_FFT_PARITY && mode != TACO_FFT_WINDOW) throw Error(TACO_ERR_INVALID, "unknown FFT mode");
    g_fft_mode = mode;
}

static int g_phase_mode = TACO_PHASE_LIBM;

int phase_mode() { return g_phase_mode; }

void set_phase_mode(int mode) {
    if (mode != TACO_PHASE_LIBM && mode != TACO_PHASE_CR) throw Error(TACO_ERR_INVALID, "unknown phase mode");
    g_phase_mode = mode;
}

double exp_in_mode(double x) {
    return g_phase_mode == TACO_PHASE_LIBM ? std::exp(x) : taco_cr::exp_cr(x, exp_table());
}

cd expi_in_mode(double x) {
    if (g_phase_mode == TACO_PHASE_LIBM) return std::exp(cd{0.0, x});
    double s, c;
    taco_cr::sincos_cr(x, &s, &c, cr_table());
    return cd(c, s);
}

void libm_phase_table(const double* omega, std::size_t nrows, std::size_t len, double dt, double* out) {
    constexpr std::size_t kPiece = 1 << 14;
    const std::size_t pieces = (len + kPiece - 1) / kPiece;
    const long total = static_cast<long>(nrows * pieces);
#pragma omp parallel for schedule(dynamic, 4)
    for (long w = 0; w < total; ++w) {
        const std::size_t r = static_cast<std::size_t>(w) / pieces, k0 = (static_cast<std::size_t>(w) % pieces) * kPiece;
        const std::size_t k1 = std::min(len, k0 + kPiece);
        const double om = omega[r];
        double* row = out + 2 * r * len;
        for (std::size_t k = k0; k < k1; ++k) {
            const double tk = static_cast<double>(k) * dt;  // tcl4_kernels.cpp:116-120
            const cd z = std::exp(cd{0.0, om * tk});
            row[2 * k] = z.real();
            row[2 * k + 1] = z.imag();
        }
    }
}

// The exponentials of the bath correlation (J's Ohmic factor, the KMS
// factors) and the trapz phase step follow the phase mode: the platform
// libm's std::exp, as the reference calls, or the correctly rounded values
// (cr_exp.hpp, cr_sincos.hpp) that the device producer computes with the
// same code.
static double spectral_J(int kind, double p0, double p1, double w) {
    if (!(w > 0.0)) return 0.0;
    if (kind == 0) return p0 * w * exp_in_mode(-w / p1);
    return 2.0 * p0 * p1 * w / (w * w + p1 * p1);
}

void bcf_fft(std::size_t N, double dt, int kind, double p0, double p1, double beta, double* C_out) {
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "bcf_fft: dt must be > 0");
    if (kind != 0 && kind != 1) throw Error(TACO_ERR_INVALID, "bcf_fft: unknown spectral density kind");
    bcf_fft_fn(N, dt, [=](double w) { return spectral_J(kind, p0, p1, w); }, beta, C_out);
}

// The same for any spectral density J (sampled at w > 0 only).
void bcf_fft_fn(std::size_t N, double dt, const std::function<double(double)>& J, double beta, double* C_out) {
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "bcf_fft: dt must be > 0");
    const std::size_t Nt = N + 1;
    std::size_t Nfft = 2 * Nt;
    if ((Nfft & (Nfft - 1)) != 0) Nfft = next_pow2(Nfft);
    const double tf = dt * static_cast<double>(Nfft);
    const double domega = 2.0 * kPi / tf;
    const double omegaNy = kPi / dt;
    const std::size_t posCount = Nfft / 2 - 1;
    std::vector<cd> S(Nfft, cd(0.0, 0.0));
    if (std::isinf(beta)) {
        for (std::size_t k = 1; k <= posCount; ++k) {
            const double w = domega * static_cast<double>(k);
            S[k] = cd((kPi / 2.0) * J(w), 0.0);
        }
        S[Nfft / 2] = cd(0.25 * kPi * J(omegaNy), 0.0);
    } else {
        const double r = exp_in_mode(-beta * domega);
        double KMS = r;
        for (std::size_t k = 1; k <= posCount; ++k) {
            const double w = domega * static_cast<double>(k);
            const double Sp = (kPi / 2.0) * J(w) / (1.0 - KMS);
            S[k] = cd(Sp, 0.0);
            S[Nfft - k] = cd(Sp * KMS, 0.0);
            KMS *= r;
        }
        const double ed = r;
        S[0] = cd(0.5 * (kPi / 2.0) * J(domega) * (1.0 + ed) / (1.0 - ed), 0.0);
        const double eNy = exp_in_mode(beta * omegaNy);
        S[Nfft / 2] = cd(0.25 * kPi * J(omegaNy) * (1.0 + 2.0 / (eNy - 1.0)), 0.0);
    }
    fft_host(S, true);
    const double scale = domega / kPi;
    for (std::size_t n = 0; n < Nt; ++n) {
        C_out[2 * n] = S[n].real() * scale;
        C_out[2 * n + 1] = S[n].imag() * scale;
    }
}

void gamma_trapz(const double* Cin, std::size_t N, double dt, const double* omegas, int M, double* Gout) {
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "gamma_trapz: dt must be > 0");
    if (N == 0 || M <= 0) return;
    const cd* C = reinterpret_cast<const cd*>(Cin);
    cd* G = reinterpret_cast<cd*>(Gout);
    const cd half_dt(dt / 2.0, 0.0);
    for (int j = 0; j < M; ++j) {
        cd* col = G + static_cast<std::size_t>(j) * N;
        col[0] = cd(0.0, 0.0);
        const cd step = expi_in_mode(omegas[j] * dt);  // gamma.hpp:185
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

// Hermitian eigensystem (system_prep.hpp: cyclic complex Jacobi, eigenvalues
// ascending, gauge fixed), the same steps the device builder runs.
static void eigh(const cd* Hin, int n, double* w, cd* Vout) {
    using namespace taco_sys;
    const std::size_t N = static_cast<std::size_t>(n);
    std::vector<Cx> A(N * N), V(N * N, cx(0.0, 0.0)), col(N);
    for (std::size_t e = 0; e < N * N; ++e) A[e] = cx(Hin[e].real(), Hin[e].imag());
    for (std::size_t r = 0; r < N; ++r) {
        if (std::abs(A[r + r * N].im) > 1e-12 * (1.0 + std::abs(A[r + r * N].re)))
            throw Error(TACO_ERR_INVALID, "Eigensystem: Hamiltonian must be Hermitian");
        V[r + r * N] = cx(1.0, 0.0);
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        if (converged(A.data(), n)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const Rot g = rotation(A.data(), n, p, q);
                if (g.skip) continue;
                for (int r = 0; r < n; ++r) rot_cols(A.data(), n, p, q, g, r);
                for (int c = 0; c < n; ++c) rot_rows(A.data(), n, p, q, g, c);
                rot_fix(A.data(), n, p, q);
                for (int r = 0; r < n; ++r) rot_cols(V.data(), n, p, q, g, r);
            }
    }
    std::vector<int> order(N);
    sort_diag(A.data(), n, order.data());
    for (std::size_t c = 0; c < N; ++c) {
        const int src = order[c];
        w[c] = A[static_cast<std::size_t>(src) * (N + 1)].re;
        gauge_column(V.data(), n, src, col.data());
        for (std::size_t r = 0; r < N; ++r) Vout[r + c * N] = cd(col[r].re, col[r].im);
    }
}

void system_build(const double* Hin, int N, const double* A_lab, int num_ops, double freq_tol, double* eps,
                  double* Uout, double* A_eig, int* nf_out, double* omegas, int* mirror, int* p2f,
                  int* zero_index) {
    if (N <= 0) throw Error(TACO_ERR_INVALID, "Eigensystem: Hamiltonian dimension must be > 0");
    if (num_ops < 0) throw Error(TACO_ERR_INVALID, "system_build: num_ops must be >= 0");
    const std::size_t NN = static_cast<std::size_t>(N);
    const cd* H = reinterpret_cast<const cd*>(Hin);
    for (std::size_t r = 0; r < NN; ++r)
        for (std::size_t c = 0; c < NN; ++c)
            if (std::abs(H[r + c * NN] - std::conj(H[c + r * NN])) > 1e-12 * (1.0 + std::abs(H[r + c * NN])))
                throw Error(TACO_ERR_INVALID, "Eigensystem: Hamiltonian must be Hermitian");
    cd* U = reinterpret_cast<cd*>(Uout);
    eigh(H, N, eps, U);
    // build_frequency_buckets (system.hpp:138-157), mirror and zero bucket
    const double tol = std::max(freq_tol, std::numeric_limits<double>::epsilon());
    std::vector<double> rep(NN * NN);
    std::vector<int> idx(NN * NN), perm(NN * NN), inv(NN * NN);
    *nf_out = taco_sys::buckets(eps, N, tol, rep.data(), idx.data(), perm.data(), inv.data(), omegas, mirror, p2f,
                                zero_index);
    // A_eig = U^H A U
    using taco_sys::Cx;
    std::vector<Cx> Uc(NN * NN), Ac(NN * NN), T(NN * NN);
    for (std::size_t e = 0; e < NN * NN; ++e) Uc[e] = taco_sys::cx(U[e].real(), U[e].imag());
    const cd* Al = reinterpret_cast<const cd*>(A_lab);
    cd* Ae = reinterpret_cast<cd*>(A_eig);
    for (int o = 0; o < num_ops; ++o) {
        const cd* A = Al + static_cast<std::size_t>(o) * NN * NN;
        for (std::size_t e = 0; e < NN * NN; ++e) Ac[e] = taco_sys::cx(A[e].real(), A[e].imag());
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) T[r + c * NN] = taco_sys::uha(Uc.data(), Ac.data(), N, r, c);
        cd* E = Ae + static_cast<std::size_t>(o) * NN * NN;
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) {
                const Cx v = taco_sys::tu(T.data(), Uc.data(), N, r, c);
                E[r + c * NN] = cd(v.re, v.im);
            }
    }
}

// sin(j/64), cos(j/64), j < kTabN, as double-doubles from quad precision.
const double* cr_table() {
    static const std::vector<double> tab = [] {
        std::vector<double> t(4 * taco_cr::kTabN);
        for (int j = 0; j < taco_cr::kTabN; ++j) {
            const __float128 a = static_cast<__float128>(j) / 64;
            const __float128 sv = sinq(a), cv = cosq(a);
            const double sh = static_cast<double>(sv), ch = static_cast<double>(cv);
            t[4 * j] = sh;
            t[4 * j + 1] = static_cast<double>(sv - sh);
            t[4 * j + 2] = ch;
            t[4 * j + 3] = static_cast<double>(cv - ch);
        }
        return t;
    }();
    return tab.data();
}

// exp(j/64) (j = -23..23) and 1/n! (n <= 12) as double-doubles from quad
// precision, and ln2 in three parts.
const double* exp_table() {
    static const std::vector<double> tab = [] {
        std::vector<double> t(taco_cr::kExpTabSize);
        auto put = [&](int at, __float128 v) {
            const double h = static_cast<double>(v);
            t[at] = h;
            t[at + 1] = static_cast<double>(v - h);
        };
        for (int j = -taco_cr::kExpJ; j <= taco_cr::kExpJ; ++j)
            put(2 * (j + taco_cr::kExpJ), expq(static_cast<__float128>(j) / 64));
        __float128 f = 1;
        for (int n = 0; n <= taco_cr::kExpDeg; ++n) {
            if (n > 0) f *= n;
            put(taco_cr::kExpTabInv + 2 * n, 1 / f);
        }
        // ln2 = L0 + L1 + L2, each the double nearest the remainder (from an
        // 80-digit evaluation)
        t[taco_cr::kExpTabLn2] = 0x1.62e42fefa39efp-1;
        t[taco_cr::kExpTabLn2 + 1] = 0x1.abc9e3b39803fp-56;
        t[taco_cr::kExpTabLn2 + 2] = 0x1.7b57a079a1934p-111;
        return t;
    }();
    return tab.data();
}

void cr_exp_host(const double* x, std::size_t n, double* out) {
    const double* tab = exp_table();
    for (std::size_t k = 0; k < n; ++k) out[k] = taco_cr::exp_cr(x[k], tab);
}

void cr_sincos_host(const double* x, std::size_t n, double* s, double* c) {
    const double* tab = cr_table();
    for (std::size_t k = 0; k < n; ++k) taco_cr::sincos_cr(x[k], s + k, c + k, tab);
}

}  // namespace taco_engine

namespace taco_engine {

// std::stable_sort of order by key[order[.]] on all host cores: chunks sorted
// in parallel, then merged pairwise (inplace_merge keeps equal keys of the
// left run first), so the result is the unique stable order.
void triple_omega_keys(const double* omegas, int nf, std::uint64_t* key) {
    const long nf3 = static_cast<long>(nf) * nf * nf;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < nf3; ++t) {
        const long k = t % nf, j = (t / nf) % nf, i = t / (static_cast<long>(nf) * nf);
        const double om = (omegas[i] + omegas[j]) + omegas[k];
        std::memcpy(&key[t], &om, 8);
    }
}

void stable_sort_by_key(std::vector<int>& order, const std::uint64_t* key) {
    const long n = static_cast<long>(order.size());
    auto less = [key](int x, int y) { return key[x] < key[y]; };
    int P = 1;
#ifdef _OPENMP
    P = std::min(16, omp_get_max_threads());
#endif
    if (n < 65536 || P < 2) {
        std::stable_sort(order.begin(), order.end(), less);
        return;
    }
    std::vector<long> b(static_cast<std::size_t>(P) + 1);
    for (int q = 0; q <= P; ++q) b[static_cast<std::size_t>(q)] = n * q / P;
#pragma omp parallel for schedule(static, 1) num_threads(P)
    for (int q = 0; q < P; ++q)
        std::stable_sort(order.begin() + b[static_cast<std::size_t>(q)], order.begin() + b[static_cast<std::size_t>(q) + 1],
                         less);
    for (int w = 1; w < P; w *= 2) {
#pragma omp parallel for schedule(static, 1) num_threads(P)
        for (int q = 0; q < P; q += 2 * w) {
            if (q + w >= P) continue;
            const long lo = b[static_cast<std::size_t>(q)], mid = b[static_cast<std::size_t>(q + w)],
                       hi = b[static_cast<std::size_t>(std::min(q + 2 * w, P))];
            std::inplace_merge(order.begin() + lo, order.begin() + mid, order.begin() + hi, less);
        }
    }
}

}  // namespace taco_engine
