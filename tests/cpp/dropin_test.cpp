// Drop-in check of include/taco/tcl4_hip.hpp: the reference-style C++ API
// bound to minimal stand-ins of Eigen::MatrixXcd / sys::System / Exec.
//   dropin_test nogpu   -> the builders raise std::runtime_error (no device)
//   dropin_test gpu     -> spin-boson L4 at t = 0, 7, 100 printed with 17
//                          digits, the exception types of bad inputs, then
//                          every reference-shaped entry point called with the
//                          reference's argument lists: compute_triple_kernels
//                          (and its _cuda name), the propagate_rk4_dense_serial
//                          / _omp overloads, rk4_update_serial, and on device
//                          pointers build_mikx_device,
//                          assemble_liouvillian_cuda_device[_raw] and
//                          compute_fcr_convolution_batched
#define TACO_HIP_PROVIDE_CUDA_NAMES
#include <hip/hip_runtime_api.h>

#include <cmath>
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
// the Eigen-like pair: Mat's column vector is Vec (Eigen: MatrixXcd -> VectorXcd)
struct MatV : Mat {
    using ColumnVector = Vec;
    using Mat::Mat;
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

static void hip_ok(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}
template <class T>
static T* to_device(const T* host, std::size_t n) {
    void* p = nullptr;
    hip_ok(hipMalloc(&p, n * sizeof(T)));
    hip_ok(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    return static_cast<T*>(p);
}
template <class T>
static std::vector<T> to_host(const T* dev, std::size_t n) {
    std::vector<T> h(n);
    hip_ok(hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

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

    // compute_triple_kernels(system, gamma_series, dt, nmax, method, exec), no template arguments
    MatV Gv(static_cast<int>(Nt), nf);
    Gv.v = G.v;
    auto series = taco::tcl4::compute_triple_kernels(S, Gv, dt, 2, FCRMethod::Convolution, ex);
    auto series_cuda = taco::tcl4::compute_triple_kernels_cuda(S, Gv, dt, 2, FCRMethod::Convolution, ex);
    static_assert(std::is_same_v<decltype(series.F[0][0][0]), Vec&>, "series of the matrix's column vector");
    std::printf("series_len %d\n", series.F[1][2][0].size());
    for (int t : {1, 5, 300})
        std::printf("fcr %d %.17g %.17g %.17g %.17g %.17g %.17g\n", t, series.F[1][2][0].data()[t].real(),
                    series.F[1][2][0].data()[t].imag(), series.C[0][1][2].data()[t].real(),
                    series.C[0][1][2].data()[t].imag(), series.R[2][2][1].data()[t].real(),
                    series.R[2][2][1].data()[t].imag());
    std::printf("series_cuda_equal %d\n", series_cuda.R[2][2][1].v == series.R[2][2][1].v ? 1 : 0);

    // RK4 overloads (rk4_dense.hpp:85-129) and the step form (:40-45)
    auto one = [](std::vector<Mat>& v) { for (auto& m : v) for (int d = 0; d < 4; ++d) m(d, d) = cd(-0.1 * (d + 1), 0.0); };
    std::vector<Mat> Lh(10, Mat(4, 4));
    one(Lh);
    Vec r1(4), r2(4), r3(4), r4(4);
    for (auto* rv : {&r1, &r2, &r3, &r4}) for (int d = 0; d < 4; ++d) rv->data()[d] = 1.0;
    int ns1 = 0;
    taco::tcl::propagate_rk4_dense_serial(Ls, r1, 0.0, 0.1, [&](double, const Vec&) { ++ns1; }, 5);
    taco::tcl::propagate_rk4_dense_serial(Ls, Lh, r2, 0.0, 0.1);
    taco::tcl::propagate_rk4_dense_omp(Ls, r3, 0.0, 0.1);
    taco::tcl::propagate_rk4_dense_omp(Ls, Lh, r4, 0.0, 0.1, std::function<void(double, const Vec&)>{}, 2);
    std::printf("rk4_overloads %.17g %.17g %.17g %.17g %d\n", r1.data()[3].real(), r2.data()[3].real(),
                r3.data()[3].real(), r4.data()[3].real(), ns1);
    taco::hip::Rk4DenseWorkspaceT<Vec> ws;
    Vec rs(4);
    for (int d = 0; d < 4; ++d) rs.data()[d] = 1.0;
    taco::tcl::rk4_update_serial(Ls[0], Lh[0], Ls[1], rs, ws, 0.1);
    taco::tcl::rk4_update_omp(Ls[0], Lh[0], Ls[1], rs, ws, 0.1);
    std::printf("rk4_update %.17g %.17g\n", rs.data()[0].real(), rs.data()[3].real());

    // device-level helpers on device pointers, slices at t = 300
    const int tsel = 300;
    const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf, N4 = 16, N6 = 64;
    std::vector<cd> Fh(nf3), Ch(nf3), Rh(nf3);
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nf; ++j)
            for (int k = 0; k < nf; ++k) {
                const std::size_t e = (static_cast<std::size_t>(i) * nf + j) * nf + k;
                Fh[e] = series.F[i][j][k].data()[tsel];
                Ch[e] = series.C[i][j][k].data()[tsel];
                Rh[e] = series.R[i][j][k].data()[tsel];
            }
    cd *dF = to_device(Fh.data(), nf3), *dC = to_device(Ch.data(), nf3), *dR = to_device(Rh.data(), nf3);
    int* dpf = to_device(pf.data(), static_cast<std::size_t>(N * N));
    cd* dops = to_device(S.A_eig[0].data(), static_cast<std::size_t>(N * N));
    std::vector<cd> zeros(N6);
    cd *dM = to_device(zeros.data(), N4), *dI = to_device(zeros.data(), N4), *dK = to_device(zeros.data(), N4),
       *dX = to_device(zeros.data(), N6), *dGW = to_device(zeros.data(), N4), *dGWr = to_device(zeros.data(), N4);
    hipStream_t st = nullptr;
    hip_ok(hipStreamCreate(&st));
    taco::tcl4::cuda_mikx::build_mikx_device(taco::tcl4::cuda_mikx::MikxDeviceInputs{dF, dC, dR, dpf, N, nf},
                                             taco::tcl4::cuda_mikx::MikxDeviceOutputs{dM, dI, dK, dX}, st);
    taco::tcl4::assemble_liouvillian_cuda_device(dM, dI, dK, dX, dops, N, 1, dGW, st);
    taco::tcl4::assemble_liouvillian_cuda_device_raw(dM, dI, dK, dX, dops, N, 1, dGWr, st);
    hip_ok(hipStreamSynchronize(st));
    const std::vector<cd> GW = to_host(dGW, N4), GWr = to_host(dGWr, N4);
    std::vector<cd> L4s(N4), GWs(N4);
    check(taco_l4_from_slices(taco::hip::context(0), reinterpret_cast<const double*>(Fh.data()),
                              reinterpret_cast<const double*>(Ch.data()), reinterpret_cast<const double*>(Rh.data()),
                              pf.data(), N, nf, reinterpret_cast<const double*>(S.A_eig[0].data()), 1,
                              reinterpret_cast<double*>(L4s.data()), reinterpret_cast<double*>(GWs.data())));
    double dgw = 0.0, gmax = 0.0, dsym = 0.0;
    for (int rr = 0; rr < 4; ++rr)
        for (int c = 0; c < 4; ++c) {
            const std::size_t e = static_cast<std::size_t>(rr + 4 * c), et = static_cast<std::size_t>(c + 4 * rr);
            dgw = std::max(dgw, std::abs(GW[e] - GWs[e]));
            gmax = std::max(gmax, std::abs(GWs[e]));
            dsym = std::max(dsym, std::abs(GW[e] - (GWr[e] + std::conj(GWr[et]))));
        }
    std::printf("device_gw rel %.3e sym %.3e\n", dgw / gmax, dsym / gmax);
    const std::vector<cd> Mh = to_host(dM, N4), Xh = to_host(dX, N6);
    std::printf("mikx M0 %.17g %.17g X5 %.17g %.17g\n", Mh[0].real(), Mh[0].imag(), Xh[5].real(), Xh[5].imag());

    // compute_fcr_convolution_batched: F/C/R of (1, 2, k), k = 0..2, [Nt, batch]
    std::vector<int> mih(mi.begin(), mi.begin() + nf);
    cd* dG = to_device(G.data(), Nt * static_cast<std::size_t>(nf));
    double* dom = to_device(om.data(), static_cast<std::size_t>(nf));
    int* dmi = to_device(mih.data(), static_cast<std::size_t>(nf));
    std::vector<cd> zt(Nt * 3);
    cd *bF = to_device(zt.data(), zt.size()), *bC = to_device(zt.data(), zt.size()), *bR = to_device(zt.data(), zt.size());
    taco::tcl4::cuda_fcr::FcrWorkspace fws;
    taco::tcl4::cuda_fcr::compute_fcr_convolution_batched(
        taco::tcl4::cuda_fcr::FcrDeviceInputs{dG, dom, dmi, Nt, static_cast<std::size_t>(nf), dt},
        taco::tcl4::cuda_fcr::FcrBatch{3, 0, 1, 2, 0, bF, bC, bR}, fws, st);
    const std::vector<cd> hF = to_host(bF, zt.size()), hC = to_host(bC, zt.size()), hR = to_host(bR, zt.size());
    int same = 1;
    for (int b = 0; b < 3; ++b)
        for (std::size_t t = 0; t < Nt; ++t) {
            same &= hF[t + Nt * b] == series.F[1][2][b].data()[t];
            same &= hC[t + Nt * b] == series.C[1][2][b].data()[t];
            same &= hR[t + Nt * b] == series.R[1][2][b].data()[t];
        }
    std::printf("fcr_batched_equal %d\n", same);
    for (void* p : {static_cast<void*>(dF), static_cast<void*>(dC), static_cast<void*>(dR), static_cast<void*>(dpf),
                    static_cast<void*>(dops), static_cast<void*>(dM), static_cast<void*>(dI), static_cast<void*>(dK),
                    static_cast<void*>(dX), static_cast<void*>(dGW), static_cast<void*>(dGWr), static_cast<void*>(dG),
                    static_cast<void*>(dom), static_cast<void*>(dmi), static_cast<void*>(bF), static_cast<void*>(bC),
                    static_cast<void*>(bR)})
        hip_ok(hipFree(p));
    hip_ok(hipStreamDestroy(st));
    return bad == 0 ? 0 : 1;
}
