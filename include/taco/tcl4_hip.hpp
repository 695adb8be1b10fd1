// tcl4_hip.hpp — C++ drop-in for the reference's GPU TCL4 builders, on the
// HIP engine's C ABI (taco_tcl4.h).
//
// The functions keep the reference signatures and semantics:
//   build_TCL4_generator_cuda_fused        cpp/include/taco/backend/cuda/tcl4_fused_cuda.hpp:16-21
//   build_TCL4_generator_cuda_fused_batch  tcl4_fused_cuda.hpp:25-41 (with the cuda_fcr_ms overload)
//   build_correction_series_cuda_fused     tcl4_fused_cuda.hpp:44-48
//   compute_triple_kernels[_cuda]          cpp/include/taco/tcl4.hpp:37-42,
//                                          cpp/include/taco/backend/cuda/tcl4_kernels_cuda.hpp:7-12
//   propagate_rk4_dense_serial/_omp        cpp/include/taco/rk4_dense.hpp:85-129 (L_series overloads)
//   rk4_update_serial/_omp                 rk4_dense.hpp:40-45, 54-59 (L0, Lhalf, L1 step form)
//   build_mikx_device                      cpp/include/taco/backend/cuda/tcl4_mikx_cuda.hpp:44-46
//   assemble_liouvillian_cuda_device[_raw] cpp/include/taco/backend/cuda/tcl4_assemble_cuda.hpp:24-44
//   compute_fcr_convolution_batched        cpp/include/taco/backend/cuda/tcl4_fcr_kernels_cuda.hpp:88-91
// under the names *_hip_* and, when TACO_HIP_PROVIDE_CUDA_NAMES is defined,
// also under the reference's *_cuda_* names so existing callers relink
// unchanged.  They are templates over the caller's types: with the reference
// headers they bind to taco::sys::System, Eigen::MatrixXcd/VectorXcd and
// taco::Exec; any types exposing the same members work as well:
//   System: eig.dim, eig.eps (indexable), fidx.buckets[b].omega and .pairs,
//           fidx.tol, A_eig (vector of matrices)
//   Matrix: rows(), cols(), data() (column-major std::complex<double>),
//           constructor Matrix(rows, cols)
//   Exec:   gpu_id
//   Vector: size(), data(), constructor Vector(n)
// compute_triple_kernels returns the reference's TripleKernelSeries when
// TACO_HIP_REFERENCE_SERIES names it (e.g. -DTACO_HIP_REFERENCE_SERIES=
// taco::tcl4::TripleKernelSeries with the reference headers included), else
// taco::hip::TripleKernelSeriesT of the gamma matrix's column-vector type.
// Errors come back as the reference's exception types (tcl4_fused_cuda.cu:
// 820-851, 1107-1216): std::invalid_argument, std::out_of_range,
// std::runtime_error, std::overflow_error.
#pragma once

#include <complex>
#include <cstddef>
#include <functional>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../taco_tcl4.h"

namespace taco::hip {

inline void throw_status(int rc) {
    if (rc == TACO_OK) return;
    const std::string msg = taco_last_error();
    switch (rc) {
        case TACO_ERR_INVALID: throw std::invalid_argument(msg);
        case TACO_ERR_RANGE: throw std::out_of_range(msg);
        case TACO_ERR_OVERFLOW: throw std::overflow_error(msg);
        default: throw std::runtime_error(msg);
    }
}

// One context per host thread and device, like the reference's thread_local
// FusedCudaWorkspace (tcl4_fused_cuda.cu:676-685).
inline taco_ctx* context(int device) {
    struct Holder {
        std::map<int, taco_ctx*> ctx;
        ~Holder() {
            for (auto& kv : ctx) taco_ctx_destroy(kv.second);
        }
    };
    thread_local Holder h;
    auto it = h.ctx.find(device);
    if (it != h.ctx.end()) return it->second;
    taco_ctx* c = nullptr;
    throw_status(taco_ctx_create(device, &c));
    h.ctx.emplace(device, c);
    return c;
}

// Flat inputs of a system: bucket omegas, mirror and pair map exactly as
// build_map (tcl4.cpp:40-101) derives them, and A_eig column-major.
struct FlatSystem {
    int N = 0, nf = 0, num_ops = 0;
    std::vector<double> omegas, eps;
    std::vector<int> mirror, pair_to_freq;
    std::vector<std::complex<double>> ops;
};

template <class System>
FlatSystem flatten(const System& system) {
    FlatSystem f;
    f.N = static_cast<int>(system.eig.dim);
    f.nf = static_cast<int>(system.fidx.buckets.size());
    f.num_ops = static_cast<int>(system.A_eig.size());
    f.omegas.resize(static_cast<std::size_t>(f.nf));
    f.pair_to_freq.assign(static_cast<std::size_t>(f.N) * f.N, -1);
    for (int b = 0; b < f.nf; ++b) {
        f.omegas[static_cast<std::size_t>(b)] = system.fidx.buckets[static_cast<std::size_t>(b)].omega;
        for (const auto& pr : system.fidx.buckets[static_cast<std::size_t>(b)].pairs)
            f.pair_to_freq[static_cast<std::size_t>(pr.first) + static_cast<std::size_t>(pr.second) * f.N] = b;
    }
    const double tol = std::max(1e-12, static_cast<double>(system.fidx.tol));
    f.mirror.resize(static_cast<std::size_t>(f.nf));
    for (int b = 0; b < f.nf; ++b) {
        const double w = f.omegas[static_cast<std::size_t>(b)];
        int best = b;
        if (std::abs(w) > tol) {
            double best_abs = std::numeric_limits<double>::infinity();
            for (int bp = 0; bp < f.nf; ++bp) {
                const double dw = std::abs(f.omegas[static_cast<std::size_t>(bp)] + w);
                if (dw < best_abs) {
                    best_abs = dw;
                    best = bp;
                }
            }
        }
        f.mirror[static_cast<std::size_t>(b)] = best;
    }
    f.eps.resize(static_cast<std::size_t>(f.N));
    for (int i = 0; i < f.N; ++i) f.eps[static_cast<std::size_t>(i)] = system.eig.eps[i];
    f.ops.resize(static_cast<std::size_t>(f.num_ops) * f.N * f.N);
    for (int o = 0; o < f.num_ops; ++o) {
        const auto& A = system.A_eig[static_cast<std::size_t>(o)];
        if (A.rows() != f.N || A.cols() != f.N)
            throw std::invalid_argument("build_TCL4_generator_hip_fused: coupling operator has wrong shape");
        const std::complex<double>* d = A.data();
        for (int e = 0; e < f.N * f.N; ++e) f.ops[static_cast<std::size_t>(o) * f.N * f.N + e] = d[e];
    }
    return f;
}

inline const double* dp(const std::complex<double>* p) { return reinterpret_cast<const double*>(p); }
inline double* dp(std::complex<double>* p) { return reinterpret_cast<double*>(p); }

// Column-vector type of a matrix type: M::ColumnVector when it declares one,
// Eigen's Matrix<Scalar, Dynamic, 1> for Eigen matrices, else a std::vector.
template <class M, class = void>
struct column_vector {
    using type = std::vector<std::complex<double>>;
};
template <class M>
struct column_vector<M, std::void_t<typename M::ColumnVector>> {
    using type = typename M::ColumnVector;
};
#ifdef EIGEN_WORLD_VERSION
template <class S, int R, int C, int O, int MR, int MC>
struct column_vector<Eigen::Matrix<S, R, C, O, MR, MC>, void> {
    using type = Eigen::Matrix<S, Eigen::Dynamic, 1>;
};
#endif

// TripleKernelSeries (tcl4.hpp:30-34) over a vector type.
template <class Vector>
struct TripleKernelSeriesT {
    std::vector<std::vector<std::vector<Vector>>> F, C, R;
};

template <class Matrix>
struct series_of {
#ifdef TACO_HIP_REFERENCE_SERIES
    using type = TACO_HIP_REFERENCE_SERIES;
#else
    using type = TripleKernelSeriesT<typename column_vector<Matrix>::type>;
#endif
};

// Reference-style workspace of the step form (rk4_dense.hpp:13-21); the GPU
// step keeps its intermediates on the device, so it is only resized.
template <class Vector>
struct Rk4DenseWorkspaceT {
    Vector k1, k2, k3, k4, tmp;
    void resize(long n) { k1 = k2 = k3 = k4 = tmp = Vector(static_cast<int>(n)); }
};

}  // namespace taco::hip

namespace taco::tcl4 {

// Batched fused HIP path (tcl4_fused_cuda.hpp:25-41).  Empty time_indices
// means every t < Nt.  fcr_ms (nullable) receives the F/C/R stage's device time.
template <class System, class Matrix, class FCRMethodT, class ExecT>
std::vector<Matrix> build_TCL4_generator_hip_fused_batch(const System& system, const Matrix& gamma_series, double dt,
                                                         const std::vector<std::size_t>& time_indices,
                                                         FCRMethodT method, const ExecT& exec,
                                                         double* fcr_ms = nullptr) {
    if (static_cast<int>(method) != TACO_FCR_CONVOLUTION)  // tcl4_fused_cuda.cu:820-822
        throw std::invalid_argument("build_TCL4_generator_cuda_fused: only FCRMethod::Convolution is supported");
    const auto f = taco::hip::flatten(system);
    const std::size_t Nt = static_cast<std::size_t>(gamma_series.rows());
    if (static_cast<int>(gamma_series.cols()) != f.nf)
        throw std::invalid_argument(
            "build_TCL4_generator_hip_fused_batch: gamma_series column count does not match frequency buckets");
    const std::size_t ntau = time_indices.empty() ? Nt : time_indices.size();
    const std::size_t N2 = static_cast<std::size_t>(f.N) * f.N;
    std::vector<std::complex<double>> out(ntau * N2 * N2);
    taco_ctx* ctx = taco::hip::context(exec.gpu_id);
    taco::hip::throw_status(taco_l4_batch(ctx, taco::hip::dp(gamma_series.data()), Nt, f.nf, f.omegas.data(),
                                          f.mirror.data(), f.pair_to_freq.data(), f.N,
                                          taco::hip::dp(f.ops.data()), f.num_ops, dt, static_cast<int>(method),
                                          time_indices.empty() ? nullptr : time_indices.data(),
                                          time_indices.size(), taco::hip::dp(out.data()), fcr_ms));
    std::vector<Matrix> res;
    res.reserve(ntau);
    for (std::size_t q = 0; q < ntau; ++q) {
        Matrix L(static_cast<int>(N2), static_cast<int>(N2));
        std::complex<double>* d = L.data();
        for (std::size_t e = 0; e < N2 * N2; ++e) d[e] = out[q * N2 * N2 + e];
        res.push_back(std::move(L));
    }
    return res;
}

// Single time index (tcl4_fused_cuda.hpp:16-21).
template <class System, class Matrix, class FCRMethodT, class ExecT>
Matrix build_TCL4_generator_hip_fused(const System& system, const Matrix& gamma_series, double dt,
                                      std::size_t time_index, FCRMethodT method, const ExecT& exec) {
    if (time_index >= static_cast<std::size_t>(gamma_series.rows()))
        throw std::out_of_range("build_TCL4_generator_hip_fused: time_index out of range");
    return build_TCL4_generator_hip_fused_batch(system, gamma_series, dt, std::vector<std::size_t>{time_index},
                                                method, exec)[0];
}

// Every time index (tcl4_fused_cuda.hpp:44-48).
template <class System, class Matrix, class FCRMethodT, class ExecT>
std::vector<Matrix> build_correction_series_hip_fused(const System& system, const Matrix& gamma_series, double dt,
                                                      FCRMethodT method, const ExecT& exec) {
    return build_TCL4_generator_hip_fused_batch(system, gamma_series, dt, std::vector<std::size_t>{}, method, exec);
}

// All nf^3 F/C/R series on the GPU (tcl4_kernels_cuda.hpp:7-12).  Series is
// the reference's TripleKernelSeries (F[i][j][k] vectors of length Nt) or,
// by default, taco::hip::series_of<Matrix>; Vector is its element type
// (Eigen::VectorXcd: Vector(n), data()).
template <class Series = void, class Vector = void, class System, class Matrix, class FCRMethodT, class ExecT,
          class S = std::conditional_t<std::is_void_v<Series>, typename taco::hip::series_of<Matrix>::type, Series>,
          class V = std::conditional_t<std::is_void_v<Vector>,
                                       std::decay_t<decltype(std::declval<S&>().F[0][0][0])>, Vector>>
S compute_triple_kernels_hip(const System& system, const Matrix& gamma_series, double dt, int nmax,
                             FCRMethodT method, const ExecT& exec) {
    using Vec = V;
    (void)nmax;
    const auto f = taco::hip::flatten(system);
    const std::size_t Nt = static_cast<std::size_t>(gamma_series.rows());
    if (static_cast<int>(gamma_series.cols()) != f.nf)
        throw std::invalid_argument(
            "compute_triple_kernels: gamma_series column count does not match frequency buckets");
    const std::size_t nf = static_cast<std::size_t>(f.nf), nf3 = nf * nf * nf;
    std::vector<std::complex<double>> F(nf3 * Nt), C(nf3 * Nt), R(nf3 * Nt);
    taco_ctx* ctx = taco::hip::context(exec.gpu_id);
    taco::hip::throw_status(taco_triple_kernels(ctx, taco::hip::dp(gamma_series.data()), Nt, f.nf,
                                                f.omegas.data(), f.mirror.data(), dt, static_cast<int>(method), Nt,
                                                taco::hip::dp(F.data()), taco::hip::dp(C.data()),
                                                taco::hip::dp(R.data())));
    S s;
    auto fill = [&](auto& dst, const std::vector<std::complex<double>>& src) {
        dst.resize(nf);
        for (std::size_t i = 0; i < nf; ++i) {
            dst[i].resize(nf);
            for (std::size_t j = 0; j < nf; ++j) {
                dst[i][j].resize(nf);
                for (std::size_t k = 0; k < nf; ++k) {
                    Vec v(static_cast<int>(Nt));
                    const std::size_t off = ((i * nf + j) * nf + k) * Nt;
                    for (std::size_t t = 0; t < Nt; ++t) v.data()[t] = src[off + t];
                    dst[i][j][k] = std::move(v);
                }
            }
        }
    };
    fill(s.F, F);
    fill(s.C, C);
    fill(s.R, R);
    return s;
}

// compute_triple_kernels(system, gamma_series, dt, nmax, method, exec)
// (tcl4.hpp:37-42): every backend request runs on the GPU.
template <class System, class Matrix, class FCRMethodT, class ExecT>
auto compute_triple_kernels(const System& system, const Matrix& gamma_series, double dt, int nmax, FCRMethodT method,
                            const ExecT& exec) {
    return compute_triple_kernels_hip(system, gamma_series, dt, nmax, method, exec);
}

// ---------------------------------------------------------------------------
// Device-level helpers (device pointers; the stream is a hipStream_t or NULL).
// The complex type of the pointers is any 16-byte (re, im) double pair:
// hipDoubleComplex, double2, std::complex<double>.
// ---------------------------------------------------------------------------
namespace cuda_mikx {
template <class Cplx>
struct MikxDeviceInputsT {
    const Cplx* F{nullptr};
    const Cplx* C{nullptr};
    const Cplx* R{nullptr};
    const int* pair_to_freq{nullptr};
    int N{0};
    int nf{0};
};
template <class Cplx>
struct MikxDeviceOutputsT {
    Cplx* M{nullptr};
    Cplx* I{nullptr};
    Cplx* K{nullptr};
    Cplx* X{nullptr};
};
using MikxDeviceInputs = MikxDeviceInputsT<std::complex<double>>;
using MikxDeviceOutputs = MikxDeviceOutputsT<std::complex<double>>;

// build_mikx_device(inputs, outputs, stream) (tcl4_mikx_cuda.hpp:44-46).
template <class Cplx, class Stream>
void build_mikx_device(const MikxDeviceInputsT<Cplx>& in, const MikxDeviceOutputsT<Cplx>& out, Stream stream) {
    static_assert(sizeof(Cplx) == 16, "complex type must be a (re, im) pair of doubles");
    taco::hip::throw_status(taco_mikx_device(in.F, in.C, in.R, in.pair_to_freq, in.N, in.nf, out.M, out.I, out.K,
                                             out.X, static_cast<void*>(stream)));
}
}  // namespace cuda_mikx

// assemble_liouvillian_cuda_device[_raw](dM, dI, dK, dX, d_ops, N, num_ops,
// dGW, stream) (tcl4_assemble_cuda.hpp:24-44).
template <class Cplx, class Stream>
void assemble_liouvillian_cuda_device(const Cplx* dM, const Cplx* dI, const Cplx* dK, const Cplx* dX,
                                      const Cplx* d_ops, int N, int num_ops, Cplx* dGW, Stream stream) {
    static_assert(sizeof(Cplx) == 16, "complex type must be a (re, im) pair of doubles");
    taco::hip::throw_status(taco_assemble_gw_device(dM, dI, dK, dX, d_ops, N, num_ops, dGW, 0,
                                                    static_cast<void*>(stream)));
}
template <class Cplx, class Stream>
void assemble_liouvillian_cuda_device_raw(const Cplx* dM, const Cplx* dI, const Cplx* dK, const Cplx* dX,
                                          const Cplx* d_ops, int N, int num_ops, Cplx* dGW_raw, Stream stream) {
    static_assert(sizeof(Cplx) == 16, "complex type must be a (re, im) pair of doubles");
    taco::hip::throw_status(taco_assemble_gw_device(dM, dI, dK, dX, d_ops, N, num_ops, dGW_raw, 1,
                                                    static_cast<void*>(stream)));
}

namespace cuda_fcr {
template <class Cplx>
struct FcrDeviceInputsT {
    const Cplx* gamma{nullptr};  // [Nt, nf] column-major: gamma[t + Nt*b]
    const double* omegas{nullptr};
    const int* mirror{nullptr};
    std::size_t Nt{0};
    std::size_t nf{0};
    double dt{0.0};
};
template <class Cplx>
struct FcrBatchT {
    std::size_t batch{0};
    std::size_t Nfft{0};  // informational: the engine picks Nfft by the reference's rule
    int i{0}, j{0}, k0{0};
    Cplx* F{nullptr};  // [Nt, batch]: out[t + Nt*b]
    Cplx* C{nullptr};
    Cplx* R{nullptr};
};
// The engine's FFT workspace lives in its per-thread context; this stands in
// for the reference's cuFFT/CUB workspace (tcl4_fcr_kernels_cuda.hpp:53-66).
struct FcrWorkspace {
    int gpu_id{0};
};
using FcrDeviceInputs = FcrDeviceInputsT<std::complex<double>>;
using FcrBatch = FcrBatchT<std::complex<double>>;

// compute_fcr_convolution_batched(inputs, batch, ws, stream): runs on the
// context's stream of ws.gpu_id and has completed on return (the stream
// argument is accepted for the signature; the work is ordered after earlier
// work on that context).
template <class Cplx, class Stream>
void compute_fcr_convolution_batched(const FcrDeviceInputsT<Cplx>& in, const FcrBatchT<Cplx>& b, FcrWorkspace& ws,
                                     Stream stream) {
    static_assert(sizeof(Cplx) == 16, "complex type must be a (re, im) pair of doubles");
    (void)stream;
    taco::hip::throw_status(taco_fcr_batched_dev(taco::hip::context(ws.gpu_id), in.gamma, in.omegas, in.mirror,
                                                 in.Nt, static_cast<int>(in.nf), in.dt, b.i, b.j, b.k0,
                                                 static_cast<int>(b.batch), b.F, b.C, b.R));
}
}  // namespace cuda_fcr

}  // namespace taco::tcl4

namespace taco::tcl {

// propagate_rk4_dense_serial(L_series[, L_half_series], r, t0, dt, on_sample,
// sample_every) on the GPU (rk4_dense.hpp:85-100).
template <class Matrix, class Vector>
void propagate_rk4_dense_hip(const std::vector<Matrix>& L_series, const std::vector<Matrix>* L_half_series,
                             Vector& r, double t0, double dt,
                             const std::function<void(double, const Vector&)>& on_sample = {},
                             std::size_t sample_every = 1, int gpu_id = 0) {
    if (L_series.size() < 2)
        throw std::invalid_argument("propagate_rk4_dense: L_series must have at least 2 matrices (endpoints)");
    const int D = static_cast<int>(L_series.front().rows());
    if (static_cast<int>(r.size()) != D)
        throw std::invalid_argument("propagate_rk4_dense: L_series and r dimension mismatch");
    const std::size_t steps = L_series.size() - 1, D2 = static_cast<std::size_t>(D) * D;
    std::vector<std::complex<double>> Ls((steps + 1) * D2), Lh;
    for (std::size_t i = 0; i <= steps; ++i) {
        if (L_series[i].rows() != D || L_series[i].cols() != D)
            throw std::invalid_argument("propagate_rk4_dense: L_series dimension mismatch");
        for (std::size_t e = 0; e < D2; ++e) Ls[i * D2 + e] = L_series[i].data()[e];
    }
    if (L_half_series) {
        if (L_half_series->size() != steps)
            throw std::invalid_argument(
                "propagate_rk4_dense: L_half_series must match L_series size - 1 (midpoints)");
        Lh.resize(steps * D2);
        for (std::size_t i = 0; i < steps; ++i)
            for (std::size_t e = 0; e < D2; ++e) Lh[i * D2 + e] = (*L_half_series)[i].data()[e];
    }
    std::vector<std::complex<double>> samples;
    if (on_sample) samples.resize((steps / sample_every + 2) * static_cast<std::size_t>(D));
    std::size_t ns = 0;
    taco_ctx* ctx = taco::hip::context(gpu_id);
    taco::hip::throw_status(taco_rk4_propagate(ctx, taco::hip::dp(Ls.data()),
                                               L_half_series ? taco::hip::dp(Lh.data()) : nullptr, steps, D,
                                               taco::hip::dp(r.data()), dt, sample_every,
                                               on_sample ? taco::hip::dp(samples.data()) : nullptr, &ns));
    if (on_sample) {
        Vector v(D);
        for (std::size_t q = 0; q < ns; ++q) {
            for (int e = 0; e < D; ++e) v.data()[e] = samples[q * static_cast<std::size_t>(D) + e];
            const std::size_t step = (q + 1 < ns) ? q * sample_every : steps;
            on_sample(t0 + static_cast<double>(step) * dt, v);
        }
    }
}

// The reference's overloads (rk4_dense.hpp:85-129): L_series (midpoints
// averaged) or L_series + L_half_series, serial and OpenMP twins; all run on
// GPU 0 (propagate_rk4_dense_hip takes a device id).
template <class Matrix, class Vector, class Fn = std::function<void(double, const Vector&)>>
void propagate_rk4_dense_serial(const std::vector<Matrix>& L_series, Vector& r, double t0, double dt,
                                const Fn& on_sample = {}, std::size_t sample_every = 1) {
    propagate_rk4_dense_hip<Matrix, Vector>(L_series, nullptr, r, t0, dt, on_sample, sample_every);
}
template <class Matrix, class Vector, class Fn = std::function<void(double, const Vector&)>>
void propagate_rk4_dense_serial(const std::vector<Matrix>& L_series, const std::vector<Matrix>& L_half_series,
                                Vector& r, double t0, double dt, const Fn& on_sample = {},
                                std::size_t sample_every = 1) {
    propagate_rk4_dense_hip<Matrix, Vector>(L_series, &L_half_series, r, t0, dt, on_sample, sample_every);
}
template <class Matrix, class Vector, class Fn = std::function<void(double, const Vector&)>>
void propagate_rk4_dense_omp(const std::vector<Matrix>& L_series, Vector& r, double t0, double dt,
                             const Fn& on_sample = {}, std::size_t sample_every = 1) {
    propagate_rk4_dense_hip<Matrix, Vector>(L_series, nullptr, r, t0, dt, on_sample, sample_every);
}
template <class Matrix, class Vector, class Fn = std::function<void(double, const Vector&)>>
void propagate_rk4_dense_omp(const std::vector<Matrix>& L_series, const std::vector<Matrix>& L_half_series,
                             Vector& r, double t0, double dt, const Fn& on_sample = {},
                             std::size_t sample_every = 1) {
    propagate_rk4_dense_hip<Matrix, Vector>(L_series, &L_half_series, r, t0, dt, on_sample, sample_every);
}

// One RK4 step r <- r + dt/6 (k1 + 2 k2 + 2 k3 + k4) with L at t, t + dt/2,
// t + dt (rk4_update_serial / _omp, rk4_dense.hpp:40-45, 54-59), on the GPU.
template <class Matrix, class Vector, class Workspace>
void rk4_update_serial(const Matrix& L0, const Matrix& Lhalf, const Matrix& L1, Vector& r, Workspace& ws, double dt) {
    ws.resize(static_cast<long>(r.size()));
    const std::vector<Matrix> ends{L0, L1}, mids{Lhalf};
    propagate_rk4_dense_hip<Matrix, Vector>(ends, &mids, r, 0.0, dt);
}
template <class Matrix, class Vector, class Workspace>
void rk4_update_omp(const Matrix& L0, const Matrix& Lhalf, const Matrix& L1, Vector& r, Workspace& ws, double dt) {
    rk4_update_serial(L0, Lhalf, L1, r, ws, dt);
}

}  // namespace taco::tcl

#ifdef TACO_HIP_PROVIDE_CUDA_NAMES
namespace taco::tcl4 {
template <class... A>
auto build_TCL4_generator_cuda_fused(A&&... a) -> decltype(build_TCL4_generator_hip_fused(std::forward<A>(a)...)) {
    return build_TCL4_generator_hip_fused(std::forward<A>(a)...);
}
template <class... A>
auto build_TCL4_generator_cuda_fused_batch(A&&... a)
    -> decltype(build_TCL4_generator_hip_fused_batch(std::forward<A>(a)...)) {
    return build_TCL4_generator_hip_fused_batch(std::forward<A>(a)...);
}
template <class... A>
auto build_correction_series_cuda_fused(A&&... a)
    -> decltype(build_correction_series_hip_fused(std::forward<A>(a)...)) {
    return build_correction_series_hip_fused(std::forward<A>(a)...);
}
template <class... A>
auto compute_triple_kernels_cuda(A&&... a) -> decltype(compute_triple_kernels_hip(std::forward<A>(a)...)) {
    return compute_triple_kernels_hip(std::forward<A>(a)...);
}
}  // namespace taco::tcl4
#endif
