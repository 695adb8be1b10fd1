/* taco_tcl4.h — C ABI of the MI355X (gfx950) TCL4 engine.
 *
 * Conventions shared by every entry point:
 *   - complex values are interleaved (re, im) doubles, the byte layout of
 *     std::complex<double> / Eigen::MatrixXcd;
 *   - matrices are column-major (Eigen default): A(r, c) = A[r + c*rows];
 *   - gamma is the [Nt x nf] Gamma table, column-major: gamma[t + Nt*b];
 *   - pair_to_freq is N x N column-major; ops are num_ops consecutive N x N
 *     matrices (the coupling operators in the eigenbasis, A_eig);
 *   - outputs are caller-owned; every call is synchronous unless the name
 *     ends in _dev, in which case pointers are device pointers and the call
 *     returns after the work is enqueued on the context stream and
 *     completed (it synchronises the stream before returning); a name ending
 *     in _noalloc takes device pointers and a stream and only enqueues;
 *   - every function returns TACO_OK (0) or an error code; taco_last_error()
 *     gives the message of the last failure on the calling thread.
 *
 * Error codes mirror the exception types the reference throws
 * (cpp/src/backend/cuda/tcl4_fused_cuda.cu:820-851, 1107-1216):
 *   TACO_ERR_INVALID   std::invalid_argument
 *   TACO_ERR_RANGE     std::out_of_range
 *   TACO_ERR_RUNTIME   std::runtime_error (missing buckets, device errors)
 *   TACO_ERR_OVERFLOW  std::overflow_error
 */
#ifndef TACO_TCL4_H
#define TACO_TCL4_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TACO_OK 0
#define TACO_ERR_INVALID 1
#define TACO_ERR_RANGE 2
#define TACO_ERR_RUNTIME 3
#define TACO_ERR_OVERFLOW 4

/* FCRMethod (cpp/include/taco/tcl4_kernels.hpp:9-12). */
#define TACO_FCR_CONVOLUTION 0
#define TACO_FCR_DIRECT 1

typedef struct taco_ctx taco_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* taco_last_error(void);

/* Library version string. Replaces taco::version() (cpp/src/core/version.cpp:1-7). */
const char* taco_version(void);

/* One context per host thread, bound to one GPU: it owns a stream, the
 * twiddle-table cache and a grow-only device workspace, like the reference's
 * thread_local FusedCudaWorkspace (tcl4_fused_cuda.cu:676-685). */
int taco_ctx_create(int device, taco_ctx** out);
int taco_ctx_destroy(taco_ctx* ctx);
/* The HIP stream (hipStream_t) the context enqueues on. */
void* taco_ctx_stream(taco_ctx* ctx);

/* Convolution FFT padding factor, process-global like the reference
 * (set_fcr_fft_pad_factor / get_fcr_fft_pad_factor, tcl4_kernels.cpp:179-185).
 * Nfft = next_pow2(max(2*Nt-1, factor*Nt)), at least 2. */
void taco_set_fcr_fft_pad_factor(size_t factor);
size_t taco_get_fcr_fft_pad_factor(void);

/* Phase and exponential convention, process-global like the pad factor.
 * TACO_PHASE_LIBM (default): the F/C/R phases exp(+-i*Omega*(k*dt)), the
 * trapz phase step exp(i*omega_b*dt) and the bath correlation's
 * exponentials are the platform libm's std::exp values, the reference's own
 * arithmetic (tcl4_kernels.cpp:116-120, gamma.hpp:185,
 * correlation_fft.hpp:154-214).  The host evaluates them: the F/C/R phases
 * as one table row per distinct Omega over t < Nt, cached per context (and
 * per batch system or tape) and read by the kernels; a table larger than
 * TACO_LIBM_TABLE_BYTES (default 16 GiB) is refused with TACO_ERR_INVALID.
 * TACO_PHASE_CR: correctly rounded cos/sin/exp of the same arguments,
 * evaluated on the device (no table; any size).  Returns TACO_ERR_INVALID
 * for an unknown mode. */
#define TACO_PHASE_LIBM 0
#define TACO_PHASE_CR 1
int taco_set_phase_mode(int mode);
int taco_get_phase_mode(void);

/* F/C/R FFT mode, process-global.  TACO_FFT_PARITY (default): the
 * reference's radix-2 FFT with its recurrence twiddles and its Nfft rule over
 * the whole Gamma table (tcl4_kernels.cpp:31-68; bit-identical spectra).
 * TACO_FFT_WINDOW (no reference counterpart): only the Gamma rows the
 * requested window reads (t < T = max time + 1) enter the convolutions,
 * Nfft = next_pow2(max(2T - 1, pad*T)), and the twiddles are exact
 * (each rounded once from a quad-precision evaluation).  The causal
 * convolutions at t < T are the same sums, so the results differ from the
 * parity mode only by rounding, and are closer to the exact convolution
 * (the recurrence twiddles drift by ~1e-11 at 2^19); the FFT work shrinks
 * with Nfft.  Returns TACO_ERR_INVALID for an unknown mode. */
#define TACO_FFT_PARITY 0
#define TACO_FFT_WINDOW 1
int taco_set_fft_mode(int mode);
int taco_get_fft_mode(void);

/* Bytes of F/C/R slices (3 nf^3 complex values per kept time) one pass of
 * the batch builders (taco_l4_batch[_dev], taco_l4_single) may hold; 0
 * restores the default (TACO_SLICE_BYTES, else 32 GiB).  A call whose kept
 * times need more is run as windows of consecutive kept times, each an F/C/R
 * pass up to its last time followed by the L4 of its times: the time
 * decomposition of the reference's distributed builder
 * (backend/cpu/tcl4_mpi_omp.cpp:62-156), with bit-identical results and the
 * FFT work repeated per window.  Per context. */
int taco_set_slice_budget(taco_ctx* ctx, size_t bytes);
size_t taco_get_slice_budget(const taco_ctx* ctx);

/* Per-kernel HIP-event profiler of a context (off by default).  Kernel ids:
 * 0 phase A, 1 phase B, 2 phase C, 3 F/C/R combine, 4 fused L4,
 * 5 per-bucket spectra, 6 TCL2, 7 RK4.  taco_profile_get returns the summed
 * device time, launch count and algorithmic HBM bytes (compulsory reads and
 * writes of each launch) since the last reset.  No reference counterpart:
 * it replaces the reference's one-shot TACO_CUDA_*_PROFILE timers
 * (tcl4_fcr_kernels.cu:135-138, tcl4_assemble_cuda.cu:68-71). */
int taco_profile_enable(taco_ctx* ctx, int on);
int taco_profile_reset(taco_ctx* ctx);
int taco_profile_get(taco_ctx* ctx, int kernel, double* total_ms, long* launches, double* algo_bytes);

/* Tile configuration the autotuner chose for FFT length 2^log2_nfft in this
 * context, per phase A, B, C (0: 8192 values/512 threads, 1: 4096/256,
 * 2: 4096/512, 3: 2048/256, 4: 4096/1024, 5: 2048/512; -1 if not tuned
 * yet).  TACO_FFT_TILE=<0..5> forces one.  No reference counterpart (the reference's cuFFT plans are
 * internal, tcl4_fcr_kernels.cu:602-622). */
int taco_fft_tile_choice(taco_ctx* ctx, int log2_nfft, int* cfg3);

/* Correctly rounded sin/cos of each x[i] (|x| < 2^30): the phase routine of
 * the F/C/R kernels in TACO_PHASE_CR, exp(+-i*omega*(k*dt)) = (c, +-s).  The
 * reference calls std::exp of an imaginary argument (tcl4_kernels.cpp:116-120),
 * whose last bits depend on the platform libm (TACO_PHASE_LIBM reproduces
 * those).  ctx == NULL evaluates on the host, else on the GPU. */
int taco_cr_sincos(taco_ctx* ctx, const double* x, size_t n, double* s, double* c);

/* Correctly rounded exp of each x[i]: the exponentials of the bath
 * correlation (the Ohmic factor of J(w) and the KMS factors) in
 * TACO_PHASE_CR; the reference calls std::exp (correlation_fft.hpp:154-214)
 * with platform-libm last bits, which TACO_PHASE_LIBM reproduces.
 * ctx == NULL evaluates on the host, else on the GPU. */
int taco_cr_exp(taco_ctx* ctx, const double* x, size_t n, double* out);

/* --------------------------------------------------------------------------
 * Inputs of the hot path (host).  These produce what the reference's callers
 * feed the builders: sys::System (system.hpp:201-226) plus build_map
 * (tcl4.cpp:40-101), C(t) (bcf_fft_fun, correlation_fft.hpp:154-214) and
 * the Gamma table (compute_trapz_prefix_multi_matrix, gamma.hpp:166-213).
 * ------------------------------------------------------------------------ */

/* H: N x N Hermitian (lab basis).  A_lab: num_ops x (N x N).  Outputs sized
 * for nf <= N*N: eps[N] ascending, U[N*N] eigenvectors (columns),
 * A_eig[num_ops*N*N], omegas[nf], mirror[nf], pair_to_freq[N*N]. */
int taco_system_build(const double* H, int N, const double* A_lab, int num_ops, double freq_tol,
                      double* eps, double* U, double* A_eig, int* nf, double* omegas, int* mirror,
                      int* pair_to_freq, int* zero_index);

/* taco_system_build on the GPU for nsys Hamiltonians at once (N <= 64), a
 * workgroup per system running the same steps as the host builder (cyclic
 * complex Jacobi with algebraic rotations, ascending eigenvalues, the
 * largest entry of each eigenvector real positive, the reference's bucket
 * scan, mirror and zero bucket, A_eig = U^H A U): every output equals
 * taco_system_build's bit for bit.  Host arrays, per system s at offsets
 * s*N (eps), s*N*N (H, U, omegas, mirror, pair_to_freq), s*num_ops*N*N
 * (A_lab, A_eig), s (nf, zero_index); synchronous. */
int taco_system_build_hip(taco_ctx* ctx, int nsys, const double* H, int N, const double* A_lab, int num_ops,
                          double freq_tol, double* eps, double* U, double* A_eig, int* nf, double* omegas,
                          int* mirror, int* pair_to_freq, int* zero_index);

/* Spectral density kinds for taco_bcf_fft. */
#define TACO_J_OHMIC 0          /* J = p0 * w * exp(-w / p1)       (alpha, omega_c) */
#define TACO_J_DRUDE_LORENTZ 1  /* J = 2 p0 p1 w / (w^2 + p1^2)    (lambda, gamma)  */

/* C(t_n), n = 0..N (N+1 samples); beta may be +inf (zero temperature). */
int taco_bcf_fft(size_t N, double dt, int kind, double p0, double p1, double beta, double* C_out);

/* Gamma[t + Nc*b] = trapz prefix of exp(i omega_b tau) C(tau), Gamma[0] = 0.
 * The phase step exp(i omega_b dt) and the exponentials of taco_bcf_fft
 * follow the phase mode (libm std::exp or correctly rounded). */
int taco_gamma_trapz(const double* C, size_t Nc, double dt, const double* omegas, int nf, double* gamma_out);

/* Spectral density given as an expression of w (the configuration files'
 * J_expr, expression.hpp:14-26): + - * / ^ (unary minus binds tighter than
 * * / and looser than ^; ^ groups to the right), exp log sqrt abs sin cos
 * tan step/heaviside, pow max min, the constants pi and e, and the named
 * parameters given here.  Compile errors (unknown identifier or function,
 * unbalanced parentheses, wrong argument counts, an expression that does not
 * reduce to one value) return TACO_ERR_INVALID with the message in
 * taco_last_error.  Evaluation calls the platform libm. */
typedef struct taco_spectral taco_spectral;
int taco_spectral_create(const char* source, const char* const* names, const double* values, int nparams,
                         taco_spectral** out);
int taco_spectral_eval(const taco_spectral* J, const double* w, size_t n, double* out);
int taco_spectral_destroy(taco_spectral* J);
/* taco_bcf_fft for an expression spectral density (bcf_fft_fun with a
 * callable J, correlation_fft.hpp:154-214, as tcl_driver.cpp builds C(t)). */
int taco_bcf_fft_spectral(size_t N, double dt, const taco_spectral* J, double beta, double* C_out);
/* Rectangle-rule Gamma (compute_rect_prefix_multi_matrix, gamma.hpp:127-160):
 * Gamma[k + Nc*b] = sum_{m<=k} (dt phi_m) C[m], phi_m = exp(i omega_b dt)^m;
 * the phase step follows the phase mode. */
int taco_gamma_rect(const double* C, size_t Nc, double dt, const double* omegas, int nf, double* gamma_out);

/* GPU twins of the two producers above (bcf_fft_fun, correlation_fft.hpp:154-214;
 * compute_trapz_prefix_multi_matrix, gamma.hpp:166-213), bit-identical to
 * them.  taco_bath_gamma_hip: C(t_n), n = 0..N, into C_out (optional) and,
 * when nf > 0, Gamma[t + (N+1)*b] into gamma_out, without a host round trip
 * between the two; beta may be +inf.  taco_gamma_trapz_hip: Gamma from a
 * given C.  Host buffers; synchronous. */
int taco_bath_gamma_hip(taco_ctx* ctx, size_t N, double dt, int kind, double p0, double p1, double beta,
                        const double* omegas, int nf, double* C_out, double* gamma_out);
int taco_gamma_trapz_hip(taco_ctx* ctx, const double* C, size_t Nc, double dt, const double* omegas, int nf,
                         double* gamma_out);

/* --------------------------------------------------------------------------
 * Hot path (GPU).
 * ------------------------------------------------------------------------ */

/* compute_triple_kernels (tcl4.hpp:37-42; tcl4.cpp:103-183) and
 * compute_triple_kernels_cuda (tcl4_kernels_cuda.hpp:7-12).  Writes the first
 * `keep` samples (keep <= Nt; 0 means Nt) of every F/C/R series:
 * F[(idx3 * keep) + t], idx3 = (i*nf + j)*nf + k.  The FFT length follows the
 * full Nt, so the kept samples equal the reference's full series.
 * method: TACO_FCR_CONVOLUTION (the FFT route, tcl4_kernels.cpp:105-175) or
 * TACO_FCR_DIRECT (the O(T^2) sums of tcl4_kernels.cpp:188-281, summed in
 * the reference's order; for checks and short series). */
int taco_triple_kernels(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                        const int* mirror, double dt, int method, size_t keep, double* F, double* C,
                        double* R);

/* build_TCL4_generator_cuda_fused_batch (tcl4_fused_cuda.hpp:25-41):
 * L4(t) for t in tidx[0..ntau) (ntau == 0 means every t < Nt), each an
 * N^2 x N^2 column-major block of L4_out.  fcr_ms (nullable) receives the
 * device time of the F/C/R stage, like the reference's cuda_fcr_ms.
 * method as for taco_triple_kernels (the reference's fused CUDA builders
 * accept Convolution only; its CPU build_TCL4_generator takes both, and the
 * C++ wrappers of the fused names keep the CUDA rule). */
int taco_l4_batch(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                  const int* mirror, const int* pair_to_freq, int N, const double* ops, int num_ops,
                  double dt, int method, const size_t* tidx, size_t ntau, double* L4_out, double* fcr_ms);

/* Same, with gamma/omegas/mirror/pair_to_freq/ops/L4_out as device pointers
 * (inputs already resident in HBM); tidx stays a host array. */
int taco_l4_batch_dev(taco_ctx* ctx, const void* d_gamma, size_t Nt, int nf, const void* d_omegas,
                      const void* d_mirror, const void* d_pair_to_freq, int N, const void* d_ops, int num_ops,
                      double dt, int method, const size_t* tidx, size_t ntau, void* d_L4_out,
                      double* fcr_ms);

/* build_TCL4_generator_cuda_fused (tcl4_fused_cuda.hpp:16-21): one L4(t). */
int taco_l4_single(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                   const int* mirror, const int* pair_to_freq, int N, const double* ops, int num_ops,
                   double dt, int method, size_t time_index, double* L4_out);

/* assemble_liouvillian + gw_to_liouvillian on the GPU from nf^3 F/C/R slices
 * at one time (the device helpers build_mikx_device,
 * assemble_liouvillian_cuda_device_raw, tcl4_mikx_cuda.hpp:44-46 and
 * tcl4_assemble_cuda.hpp:36-44, fused).  GW_out (nullable) gets GW. */
int taco_l4_from_slices(taco_ctx* ctx, const double* Fs, const double* Cs, const double* Rs,
                        const int* pair_to_freq, int N, int nf, const double* ops, int num_ops,
                        double* L4_out, double* GW_out);

/* The TCL2 generator L_unitary(H_eff) + L_dissipator of build_tcl2_components
 * (generator.cpp:81-177) in its general form, from channel matrices: gamma_ab
 * is column-major [Nt x (nf * num_ops^2)], column (b*num_ops + beta)*num_ops
 * + alpha holding Gamma_b(alpha, beta)(t_k) (num_ops = 1 is the gamma_series
 * layout of the other entry points).  L_out [ntau][N^4] at the times tidx
 * (ntau = 0: every t < Nt), each N^2 x N^2 column-major in the reference's
 * vec convention.  As in the reference, a (b, alpha, beta) whose
 * A_alpha,b^dag A_beta,b is the zero matrix contributes nothing. */
int taco_tcl2_generator(taco_ctx* ctx, const double* gamma_ab, size_t Nt, int nf, int num_ops, const double* eps,
                        const double* ops, const int* pair_to_freq, int N, const size_t* tidx, size_t ntau,
                        double* L_out);

/* Device-level helpers of the reference's split path, device pointers,
 * enqueued on `stream` (a hipStream_t; NULL = the default stream) without
 * synchronising; they allocate nothing.
 * taco_mikx_device: build_mikx_device (tcl4_mikx_cuda.hpp:44-46) — M, I, K
 * [N^4] (N^2 x N^2 column-major) and X [N^6] (flat6 j + N(k + N(p + N(q +
 * N(r + N s))))) from the F/C/R slices [nf^3] at one time; a pair_to_freq
 * entry outside [0, nf) makes the outputs NaN.
 * taco_assemble_gw_device: assemble_liouvillian_cuda_device (raw = 0: GW =
 * T + T^H) and assemble_liouvillian_cuda_device_raw (raw = 1: T), from M, I,
 * K, X and ops [num_ops][N*N] (tcl4_assemble_cuda.hpp:24-44). */
int taco_mikx_device(const void* Fs, const void* Cs, const void* Rs, const void* pair_to_freq, int N, int nf,
                     void* M, void* I, void* K, void* X, void* stream);
int taco_assemble_gw_device(const void* M, const void* I, const void* K, const void* X, const void* ops, int N,
                            int num_ops, void* GW, int raw, void* stream);

/* compute_fcr_convolution_batched (tcl4_fcr_kernels_cuda.hpp:88-91): F, C, R
 * of the triples (i, j, k0 + b), b < batch, for every t < Nt, into device
 * buffers [batch][Nt] (out[t + Nt*b], time contiguous); device inputs gamma
 * [nf][Nt] (column-major [Nt x nf]), omegas [nf], mirror [nf].  Runs on the
 * context's stream and workspace; synchronous like the other _dev calls. */
int taco_fcr_batched_dev(taco_ctx* ctx, const void* gamma, const void* omegas, const void* mirror, size_t Nt,
                         int nf, double dt, int i, int j, int k0, int batch, void* F, void* C, void* R);

/* Allocation-free twin of taco_l4_from_slices, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Device pointers: Fs, Cs, Rs [npos][nf^3] F/C/R slices (time-major, idx3 =
 * (i*nf + j)*nf + k), pair_to_freq [N*N] int column-major with entries in
 * [0, nf) (a time whose map has an entry outside it gets NaN L4),
 * ops [num_ops][N*N], L4_out [npos][N^4], GW_out [npos][N^4] (nullable).
 * Enqueued on `stream` (a hipStream_t; NULL = the default stream) of the
 * current device and not synchronised.  The factorised kernel's per-time
 * intermediates live in a 64 MiB file-scope device array that each GPU holds
 * until the process ends; times run in launches whose intermediates fit it,
 * and a single time that needs more than 64 MiB (N and nf with
 * (num_ops^2 (2N^2 + 2N^2 nf) + N^2 nf + N^2) * 16 bytes above it) is refused
 * with TACO_ERR_INVALID.  Calls that overlap on one GPU share the array and
 * race: serialise them (one stream, or events).  No hipMalloc, hipFree or
 * other allocation happens in this call. */
int taco_l4_from_slices_noalloc(const void* Fs, const void* Cs, const void* Rs, const void* pair_to_freq,
                                int N, int nf, const void* ops, int num_ops, size_t npos, void* L4_out,
                                void* GW_out, void* stream);

/* propagate_rk4_dense_serial(L_series[, L_half_series], r, t0, dt, on_sample,
 * sample_every) (rk4_dense.hpp:85-100; rk4_dense.cpp:246-320).  L_series:
 * (steps+1) D x D blocks; L_half nullable (then 0.5*(L_i + L_{i+1})).
 * r: in/out vec(rho) of length D.  samples_out (nullable) receives r at
 * steps 0, every, 2*every, ... and the final r; *nsamples gets the count. */
int taco_rk4_propagate(taco_ctx* ctx, const double* L_series, const double* L_half, size_t steps, int D,
                       double* r, double dt, size_t sample_every, double* samples_out, size_t* nsamples);

/* The drivers' composition (TCL4_spin_boson_example.cpp:284-384): RK4 on
 * L(t) = L_TCL2(t) + L4(t) (order 4), L_TCL2 (order 2) or -i[diag(eps), .]
 * (order 0), t = 0..steps, midpoint generator 0.5*(L_n + L_{n+1}).
 * eps: eigenvalues; ops: A_eig.  r: in/out vec(rho) (eigenbasis).
 * samples_out (nullable): (steps+1) x N^2 states. */
int taco_tcl4_propagate(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                        const int* mirror, const int* pair_to_freq, int N, const double* eps,
                        const double* ops, int num_ops, double dt, size_t steps, int order, double* r,
                        double* samples_out);

/* ---- Batched runs: many independent systems in one device pass ----------
 * Each system brings its own N, buckets, Gamma table (own Nt), dt, time-index
 * set and RK4 run; systems share nothing but the context.  The reference
 * runs such sets one system at a time, each call through
 * build_TCL4_generator_cuda_fused_batch (tcl4_fused_cuda.cu:1086-1522) and a
 * separate propagate_rk4_dense (rk4_dense.cpp:246-320); a batch computes the
 * same values per system (same kernels, same arithmetic) with every input,
 * job list and output resident in HBM, no host round trip between systems,
 * the TCL4 F/C/R work done once for the union of the L4 times and the RK4
 * window, and one RK4 launch stepping every system. */
typedef struct taco_batch_system {
    const double* gamma;        /* [Nt x nf] Gamma table, column-major */
    size_t Nt;
    int nf;
    const double* omegas;       /* [nf] */
    const int* mirror;          /* [nf] */
    const int* pair_to_freq;    /* N x N column-major */
    int N;
    const double* ops;          /* num_ops N x N blocks (A_eig) */
    int num_ops;
    double dt;
    const size_t* tidx;         /* L4 output times (may be NULL when ntau = 0) */
    size_t ntau;
    size_t steps;               /* RK4 steps (0: no propagation) */
    int order;                  /* generator of the RK4 run: 4 (TCL2 + L4), 2, 0 */
    const double* eps;          /* [N] eigenvalues (needed when steps > 0) */
    const double* r0;           /* [N^2] initial vec(rho), eigenbasis (needed when steps > 0) */
} taco_batch_system;

typedef struct taco_batch taco_batch;

/* Validates and uploads every system, prepares the job lists and sizes the
 * workspace.  The host arrays are only read during this call. */
int taco_batch_create(taco_ctx* ctx, const taco_batch_system* systems, int nsys, taco_batch** out);
/* Runs the whole batch and synchronises.  use_graph = 1 captures the batch
 * into a hipGraph on the first call after an eager run (which also tunes the
 * FFT tiles) and replays it afterwards; results are identical either way. */
int taco_batch_run(taco_batch* b, int use_graph);
/* System s's L4 at its tidx, request order: ntau blocks of N^2 x N^2. */
int taco_batch_l4(taco_batch* b, int s, double* L4_out);
/* System s's RK4 states at steps 0..steps ((steps+1) x N^2, nullable) and the
 * final state (N^2, nullable). */
int taco_batch_rho(taco_batch* b, int s, double* samples_out, double* r_out);
int taco_batch_destroy(taco_batch* b);
/* One-shot create + eager run + copy-out + destroy.  L4_out[s] (ntau x N^4)
 * and rho_out[s] ((steps+1) x N^2) may be NULL. */
int taco_batched_run(taco_ctx* ctx, const taco_batch_system* systems, int nsys, double* const* L4_out,
                     double* const* rho_out);

/* ---- Derivatives (no reference counterpart: the reference has none) ------
 * Reverse mode through RK4 -> L(t) = L_TCL2(t) + L4(t) -> fused L4 -> F/C/R,
 * for the composition of taco_tcl4_propagate.  taco_tape_forward runs it,
 * writes the states r_0..r_steps ((steps+1) x N^2, nullable) and keeps on the
 * device what the reverse pass reads: the F/C/R slices of every t <= steps,
 * the generators L(t) and the RK4 states every checkpoint_every steps (the
 * states in between are recomputed from those checkpoints).
 * taco_tape_adjoint takes samples_bar = dLoss/dr_n ((steps+1) x N^2) and
 * returns the cotangents of the inputs: gamma_bar [Nt x nf] (rows past steps
 * are 0), omega_bar [nf] (through the phases exp(-+i Omega t) of F/C/R),
 * ops_bar [num_ops x N^2], eps_bar [N], r0_bar [N^2]; any may be NULL.
 * Complex cotangents follow dLoss = Re sum conj(z_bar) dz (z_bar =
 * dLoss/dRe z + i dLoss/dIm z).  A tape can be reversed any number of times. */
typedef struct taco_tape taco_tape;
int taco_tape_forward(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                      const int* mirror, const int* pair_to_freq, int N, const double* eps, const double* ops,
                      int num_ops, double dt, size_t steps, int order, const double* r0, size_t checkpoint_every,
                      double* samples_out, taco_tape** out);
int taco_tape_adjoint(taco_tape* tape, const double* samples_bar, double* gamma_bar, double* omega_bar,
                      double* ops_bar, double* eps_bar, double* r0_bar);
int taco_tape_destroy(taco_tape* tape);
/* Bath-sourced tape (no reference counterpart; for gradient callers that
 * differentiate through the bath): the Gamma table of taco_tape_forward is
 * made on the device from the bath instead of read from the host -- the
 * correlation C(t_n), n < Nt, of bcf_fft_fun (correlation_fft.hpp:154-214)
 * for spectral density `kind` (TACO_J_*) with parameters (p0, p1) at inverse
 * temperature beta > 0, then its trapz Gamma
 * (compute_trapz_prefix_multi_matrix, gamma.hpp:166-213), bit-identical to
 * taco_bath_gamma_hip(Nt - 1, ...) -- and C and Gamma stay in device memory.
 * taco_tape_adjoint_bath returns bath_bar[3] = dLoss/d(p0, p1, beta) (beta
 * finite) and omega_bar including the trapz Gamma's dependence on the bucket
 * frequencies; the other outputs are those of taco_tape_adjoint (nullable).
 * taco_tape_adjoint and taco_tape_destroy accept such a tape as well. */
int taco_tape_forward_bath(taco_ctx* ctx, int kind, double p0, double p1, double beta, size_t Nt, int nf,
                           const double* omegas, const int* mirror, const int* pair_to_freq, int N, const double* eps,
                           const double* ops, int num_ops, double dt, size_t steps, int order, const double* r0,
                           size_t checkpoint_every, double* samples_out, taco_tape** out);
int taco_tape_adjoint_bath(taco_tape* tape, const double* samples_bar, double* bath_bar, double* omega_bar,
                           double* ops_bar, double* eps_bar, double* r0_bar);
/* Multi-set tape (no reference counterpart): S parameter sets of one model
 * size -- the same N and steps; Gamma, buckets, operators, eigenvalues, dt,
 * order and r0 per set -- forward and reverse together.  Each set's F/C/R,
 * L4 and TCL2 passes run in turn; the forward RK4 of all sets is one launch
 * (a workgroup per set) and so are the reverse RK4's recompute and walk.
 * Every set's samples and cotangents equal those of taco_tape_forward /
 * taco_tape_adjoint on that set alone, bit for bit.  The sets' F/C/R slices
 * stay on the device while they fit the context's slice budget
 * (taco_set_slice_budget): whole when all sets' slices fit, else packed to
 * the entries the L4 passes read (about a third of the whole at d = 16), set
 * by set in order while their running total fits; the reverse pass
 * recomputes the slices of the sets past it in one shared buffer
 * (taco_mtape_keeps_slices: 1 when every set keeps its slices).
 * samples_out / samples_bar / the cotangent arrays hold one pointer per set
 * (layouts as in taco_tape_*); an output array or entry may be NULL. */
typedef struct taco_tape_set {
    const double* gamma;        /* [Nt x nf] Gamma table, column-major */
    size_t Nt;
    int nf;
    const double* omegas;       /* [nf] */
    const int* mirror;          /* [nf] */
    const int* pair_to_freq;    /* N x N column-major */
    int N;
    const double* eps;          /* [N] */
    const double* ops;          /* num_ops N x N blocks (A_eig) */
    int num_ops;
    double dt;
    size_t steps;
    int order;                  /* 4 (TCL2 + L4), 2, 0 */
    const double* r0;           /* [N^2] */
} taco_tape_set;
typedef struct taco_mtape taco_mtape;
int taco_mtape_forward(taco_ctx* ctx, const taco_tape_set* sets, int nsets, size_t checkpoint_every,
                       double* const* samples_out, taco_mtape** out);
int taco_mtape_adjoint(taco_mtape* mt, const double* const* samples_bar, double* const* gamma_bar,
                       double* const* omega_bar, double* const* ops_bar, double* const* eps_bar,
                       double* const* r0_bar);
/* Multi-set tape whose sets take their Gamma tables from baths on the device
 * (as taco_tape_forward_bath; sets[s].gamma is not read, sets[s].Nt is the
 * grid): bath_bar[s][3] = dLoss/d(p0, p1, beta) of set s; every set's
 * samples and cotangents equal those of taco_tape_forward_bath /
 * taco_tape_adjoint_bath on that set alone, bit for bit. */
typedef struct taco_bath_params {
    int kind;          /* TACO_J_* */
    double p0, p1;     /* spectral density parameters */
    double beta;       /* inverse temperature, > 0 */
} taco_bath_params;
int taco_mtape_forward_bath(taco_ctx* ctx, const taco_tape_set* sets, const taco_bath_params* baths, int nsets,
                            size_t checkpoint_every, double* const* samples_out, taco_mtape** out);
int taco_mtape_adjoint_bath(taco_mtape* mt, const double* const* samples_bar, double* const* bath_bar,
                            double* const* omega_bar, double* const* ops_bar, double* const* eps_bar,
                            double* const* r0_bar);
int taco_mtape_keeps_slices(const taco_mtape* mt);
/* How set s keeps its F/C/R slices: 0 recomputed in the reverse pass, 1
 * whole, 2 packed; *bytes (if not NULL) = the device bytes they take (0 when
 * recomputed).  -1 for a NULL tape or s out of range. */
int taco_mtape_set_slices(const taco_mtape* mt, size_t s, size_t* bytes);
int taco_mtape_destroy(taco_mtape* mt);
/* One-shot taco_tape_forward + taco_tape_adjoint + taco_tape_destroy. */
int taco_adjoint(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                 const int* pair_to_freq, int N, const double* eps, const double* ops, int num_ops, double dt,
                 size_t steps, int order, const double* r0, size_t checkpoint_every, const double* samples_bar,
                 double* gamma_bar, double* omega_bar, double* ops_bar, double* eps_bar, double* r0_bar);

#ifdef __cplusplus
}
#endif

#endif /* TACO_TCL4_H */
