// Drop-in check of include/taco/tcl4_hip.hpp: the reference-style C++ API
// bound to minimal stand-ins of Eigen::MatrixXcd / sys::System / Exec.
//   dropin_test nogpu   -> the builders raise std::runtime_error (no device)
//   dropin_test gpu     -> spin-boson L4 at t = 0, 7, 100 printed with 17
//                          digits, then the exception types of bad inputs
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "taco/tcl4_hip.hpp"

using cd = std::complex<double>;

struct Mat {  // column-major like Eigen::MatrixXcd
    int r = 0, c = 0;
    std::vector<cd> v;
    Mat() = default;
    Mat(int rows, int cols) : r(rows), c(cols), v(static_cast<std::size_t>(rows) * cols) {}
    int rows() const { return r; }
    int cols() const { return c; }
    cd* data() { return v.data(); }
    const cd* data() const { return v.data(); }
    cd& operator()(int i, int j) { return v[static_cast<std::size_t>(i) + static_cast<std::size_t>(j) * r]; }
};
struct Vec {
    std::vector<cd> v;
    explicit Vec(int n = 0) : v(static_cast<std::size_t>(n)) {}
    int size() const { return static_cast<int>(v.size()); }
    cd* data() { return v.data(); }
    const cd* data() const { return v.data(); }
};
struct Bucket {
    double omega;
    std::vector<std::pair<int, int>> pairs;
};
struct System {
    struct { std::size_t dim; std::vector<double> eps; } eig;
    struct { double tol = 1e-9; std::vector<Bucket> buckets; } fidx;
    std::vector<Mat> A_eig;
};
enum class FCRMethod { Convolution, Direct };
struct Exec { int gpu_id = 0; };

static void check(int rc) { taco::hip::throw_status(rc); }

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    // spin-boson: H = sz/2 + sx/2, A = sz; Ohmic alpha 0.05, wc 5, beta 1; Nt = 2048, dt = 0.005
    const int N = 2;
    Mat H(N, N), A(N, N);
    H(0, 0) = 0.5; H(1, 1) = -0.5; H(0, 1) = 0.5; H(1, 0) = 0.5;
    A(0, 0) = 1.0; A(1, 1) = -1.0;
    std::vector<double> eps(N), U(2 * N * N), Ae(2 * N * N), om(N * N);
    std::vector<int> mi(N * N), pf(N * N);
    int nf = 0, zi = 0;
    check(taco_system_build(reinterpret_cast<double*>(H.data()), N, reinterpret_cast<double*>(A.data()), 1, 1e-9,
                            eps.data(), U.data(), Ae.data(), &nf, om.data(), mi.data(), pf.data(), &zi));
    System S;
    S.eig.dim = N;
    S.eig.eps = eps;
    S.fidx.buckets.resize(static_cast<std::size_t>(nf));
    for (int b = 0; b < nf; ++b) S.fidx.buckets[static_cast<std::size_t>(b)].omega = om[static_cast<std::size_t>(b)];
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n)
            S.fidx.buckets[static_cast<std::size_t>(pf[static_cast<std::size_t>(m + n * N)])].pairs.emplace_back(m, n);
    Mat Aeig(N, N);
    std::memcpy(static_cast<void*>(Aeig.data()), Ae.data(), sizeof(cd) * N * N);
    S.A_eig.push_back(Aeig);
    const std::size_t Nt = 2048;
    const double dt = 0.005;
    std::vector<double> C(2 * Nt);
    check(taco_bcf_fft(Nt - 1, dt, TACO_J_OHMIC, 0.05, 5.0, 1.0, C.data()));
    Mat G(static_cast<int>(Nt), nf);
    check(taco_gamma_trapz(C.data(), Nt, dt, om.data(), nf, reinterpret_cast<double*>(G.data())));
    Exec ex;
    if (!gpu) {
        try {
            (void)taco::tcl4::build_TCL4_generator_hip_fused(S, G, dt, 5, FCRMethod::Convolution, ex);
        } catch (const std::runtime_error& e) {
            std::printf("nogpu: runtime_error: %s\n", e.what());
            return 0;
        }
        std::printf("nogpu: expected std::runtime_error\n");
        return 1;
    }
    auto L = taco::tcl4::build_TCL4_generator_hip_fused_batch(S, G, dt, {0, 7, 100}, FCRMethod::Convolution, ex);
    for (std::size_t q = 0; q < L.size(); ++q)
        for (int e = 0; e < 16; ++e)
            std::printf("L4 %zu %d %.17g %.17g\n", q, e, L[q].data()[e].real(), L[q].data()[e].imag());
    int bad = 0;
    try { (void)taco::tcl4::build_TCL4_generator_hip_fused(S, G, dt, Nt, FCRMethod::Convolution, ex); bad++; }
    catch (const std::out_of_range&) {}
    try { (void)taco::tcl4::build_TCL4_generator_hip_fused(S, G, dt, 3, FCRMethod::Direct, ex); bad++; }
    catch (const std::invalid_argument&) {}
    try { (void)taco::tcl4::build_TCL4_generator_hip_fused(S, G, -1.0, 3, FCRMethod::Convolution, ex); bad++; }
    catch (const std::invalid_argument&) {}
    // RK4 through the drop-in on a constant generator
    std::vector<Mat> Ls(11, Mat(4, 4));
    for (auto& m : Ls) for (int d = 0; d < 4; ++d) m(d, d) = cd(-0.1 * (d + 1), 0.0);
    Vec r(4);
    for (int d = 0; d < 4; ++d) r.data()[d] = 1.0;
    int nsamp = 0;
    taco::tcl::propagate_rk4_dense_hip<Mat, Vec>(Ls, nullptr, r, 0.0, 0.1,
                                                 [&](double, const Vec&) { ++nsamp; }, 5);
    std::printf("rk4 %.17g %d\n", r.data()[3].real(), nsamp);
    std::printf("exceptions_ok %d\n", bad == 0);
    return bad == 0 ? 0 : 1;
}
