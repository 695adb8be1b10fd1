// engine.hip — MI355X (gfx950) TCL4 hot path: F/C/R -> MIKX -> GW -> L4 and
// RK4, complex FP64, all intermediates resident in HBM.
//
// F/C/R (tcl4_kernels.cpp:105-175) need, per triple (i,j,k), the causal
// convolutions dt*ifft(fft(g_i e^{-i Omega t}) .* fft(g2))[t] with the
// reference's radix-2 recurrence-twiddle FFT of length Nfft = 2^L.  The
// engine reproduces that FFT exactly with twiddle tables generated on the
// host by the same recurrence, in three HBM passes per triple:
//
//   phase A  forward stages 1..a of the strided residue groups x[r + m 2^b]
//            (input generated on the fly: gamma column times the phase),
//   phase B  forward stages a+1..L of the groups {k = k1 mod 2^a}, the
//            spectral products with the per-bucket spectra of g2 (F) and
//            conj(g2) (C), then inverse stages 1..b of the same groups,
//   phase C  inverse stages b+1..L of the groups {t = q mod 2^b}, keeping
//            only t < T (the requested window),
//
// with a = ceil(L/2), b = L - a.  Every butterfly is the reference's
// u +/- w*x with the same w, a plain (non-fused) complex product, so the
// spectra equal the reference's bit for bit.  Groups are processed
// Stockham-style in LDS, 2048 complex values per workgroup.
//
// The per-time stage (tcl4_mikx.cpp:46-136, tcl4_assemble.cpp:29-205) is one
// fused kernel per time index that gathers M/I/K/X on the fly from the nf^3
// F/C/R slices (no N^6 X tensor), forms the NAKZWAN sum, symmetrises
// GW = T + T^H in LDS and writes the reshuffled L4 block.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "engine_internal.hpp"
#include "system_prep.hpp"

#pragma clang fp contract(off)

#include "cr_sincos.hpp"
#include "cr_exp.hpp"

namespace taco_engine {

thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

static std::size_t g_pad_factor = 0;

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw Error(TACO_ERR_RUNTIME, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// sin/cos(j/64) double-double table of the correctly rounded phase routine.
__constant__ double c_crtab[4 * taco_cr::kTabN];
// exp(j/64), 1/n! and ln2 table of the correctly rounded exp.
__constant__ double c_exptab[taco_cr::kExpTabSize];

// exp(i * arg) with correctly rounded cos/sin.
__device__ __forceinline__ double2 cexpi(double arg) {
    double s, c;
    taco_cr::sincos_cr(arg, &s, &c, c_crtab);
    return make_double2(c, s);
}

constexpr int kBlock = 256;

// FFT phases: 512-thread workgroups on a tile of 8192 complex values (G
// groups of 2^lg) in LDS.  Stages run as radix-2^r register passes: a thread
// loads the 2^r values of a closed butterfly set, performs r radix-2 stages
// on them (each butterfly exactly u +/- w*x with the stage's table
// twiddle), and writes them back, Stockham-style.  Two barriers per pass.
// Tile configurations (values per workgroup, threads); the autotuner picks
// one per FFT length and phase.  kLds leaves room for the padding slots.
template <int TILE, int NT>
struct FftCfg {
    static constexpr int kTile = TILE;
    static constexpr int kFT = NT;
    static constexpr int kPT = TILE / NT;  // values per thread (8 or 16)
    static constexpr int kLds = TILE + TILE / 8;
};
using Cfg8k = FftCfg<8192, 512>;
using Cfg4k = FftCfg<4096, 256>;
using Cfg4kW = FftCfg<4096, 512>;
using Cfg2k = FftCfg<2048, 256>;
using Cfg4kX = FftCfg<4096, 1024>;
using Cfg2kW = FftCfg<2048, 512>;
using Cfg8kX = FftCfg<8192, 1024>;
constexpr int kNumFftCfg = 7;

// Physical LDS slot of (group, position): one pad slot per 16 values and an
// odd group stride keep strided accesses off a single bank.
// G is a power of two (tile_groups halves it), so the (group, position) split
// of a group-minor element index is a mask and a shift.
struct TileLayout {
    int lg, gs, G, GS, lG;
    __device__ TileLayout(int lg_, int G_)
        : lg(lg_), gs(1 << lg_), G(G_), GS((1 << lg_) + ((1 << lg_) >> 4) + 1), lG(__ffs(G_) - 1) {}
    __device__ __forceinline__ int at(int grp, int pos) const { return grp * GS + pos + (pos >> 4); }
    __device__ __forceinline__ int grp_of(int e) const { return e & (G - 1); }
    __device__ __forceinline__ int pos_of(int e) const { return e >> lG; }
};

// Phase B with 8-value tiles runs at 4 waves per SIMD (two 512-thread or
// four 256-thread workgroups per CU, as the LDS allows): at most 128 VGPRs.
#define TACO_FFT_OCC(Cfg) __attribute__((amdgpu_waves_per_eu((Cfg::kPT <= 8 ? 4 : 1), 8)))

template <class C>
__host__ __device__ inline int tile_groups(int lg, int ngroups) {
    int G = C::kTile >> lg;
    if (G < 1) G = 1;
    if (G > ngroups) G = ngroups;
    const int GS = (1 << lg) + ((1 << lg) >> 4) + 1;
    while (G > 1 && G * GS > C::kLds) G >>= 1;
    return G;
}

// Passes of r stages; local stage sp (1-based) uses twiddle
// twp[grp*gstride + (2^(sp-1) - 1) + kk], conjugated for inverse transforms.
// Sets whose butterfly index k is kmax or above produce no value a later
// pass or the caller reads and are skipped.
template <class C, int RL, bool PRE>
__device__ __forceinline__ void tile_pass(double2* buf, const TileLayout& T, int s_lo,
                                          const double2* __restrict__ twp, long gstride, bool conj, int kmax) {
    constexpr int kFT = C::kFT;
    constexpr int R = 1 << RL;
    constexpr int SETS = C::kPT / R;
    // PRE (8 values per thread): the pass loads its sets' 2^RL - 1 twiddles
    // into registers before the tile's LDS reads, so the loads' latency
    // overlaps the reads and the barrier; otherwise (kernels with no
    // registers to spare) they are loaded as they go.
    constexpr bool kPre = PRE && RL <= 3 && C::kPT <= 8;
    const int lg = T.lg;
    const int u0bits = lg - s_lo - RL + 1;
    const int spg = 1 << (lg - RL);
    const int nsets = T.G * spg;
    double2 X[SETS][R];
    double2 W[SETS][kPre ? R - 1 : 1];
    if constexpr (kPre) {
#pragma unroll
        for (int q = 0; q < SETS; ++q) {
            const int set = threadIdx.x + q * kFT;
            const int k = (set & (spg - 1)) >> u0bits;
            if (set < nsets && k < kmax) {
                const int grp = set >> (lg - RL);
                const double2* tg = twp + grp * gstride;
#pragma unroll
                for (int i = 0; i < RL; ++i) {
                    const double2* Tw = tg + ((1 << (s_lo + i - 1)) - 1);
#pragma unroll
                    for (int j = 0; j < (1 << i); ++j) W[q][(1 << i) - 1 + j] = Tw[k + (j << (s_lo - 1))];
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < SETS; ++q) {
        const int set = threadIdx.x + q * kFT;
        const int rem = set & (spg - 1);
        const int u0 = rem & ((1 << u0bits) - 1), k = rem >> u0bits;
        if (set < nsets && k < kmax) {
            const int grp = set >> (lg - RL);
            const int pos0 = u0 + ((k * R) << u0bits);
#pragma unroll
            for (int m = 0; m < R; ++m) X[q][m] = buf[T.at(grp, pos0 + (m << u0bits))];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SETS; ++q) {
        const int set = threadIdx.x + q * kFT;
        const int rem = set & (spg - 1);
        const int u0 = rem & ((1 << u0bits) - 1), k = rem >> u0bits;
        if (set < nsets && k < kmax) {
            const int grp = set >> (lg - RL);
            const double2* tg = twp + grp * gstride;
#pragma unroll
            for (int i = 0; i < RL; ++i) {
                const int h = R >> (i + 1);
                const double2* Tw = tg + ((1 << (s_lo + i - 1)) - 1);
                double2 Y[R];
#pragma unroll
                for (int j = 0; j < (1 << i); ++j) {
                    double2 w;
                    if constexpr (kPre)
                        w = W[q][(1 << i) - 1 + j];
                    else
                        w = Tw[k + (j << (s_lo - 1))];
                    if (conj) w.y = -w.y;
#pragma unroll
                    for (int m = 0; m < h; ++m) {
                        const double2 E = X[q][m + 2 * h * j];
                        const double2 O = X[q][m + h + 2 * h * j];
                        const double2 v = cmul(w, O);
                        Y[m + h * j] = cadd(E, v);
                        Y[m + h * j + R / 2] = csub(E, v);
                    }
                }
#pragma unroll
                for (int m = 0; m < R; ++m) X[q][m] = Y[m];
            }
#pragma unroll
            for (int j = 0; j < R; ++j) buf[T.at(grp, u0 + ((k + (j << (s_lo - 1))) << u0bits))] = X[q][j];
        }
    }
    __syncthreads();
}

// All lg stages of the tile: passes of at most log2(values per thread)
// stages, stage counts as even as possible.  need: only positions below it
// are read afterwards; each pass then runs only the sets that feed them (a
// set of butterfly index k writes positions from k * 2^u0bits on and reads
// the block of R * 2^u0bits positions from k * R * 2^u0bits), which leaves
// every kept value bit-identical.
template <class C, bool PRE = false>
__device__ __forceinline__ void tile_fft(double2* buf, const TileLayout& T, const double2* __restrict__ twp, long gstride,
                                         bool conj, int need = 1 << 30) {
    const int lg = T.lg;
    if (lg <= 0) return;
    constexpr int maxr = C::kPT >= 16 ? 4 : (C::kPT >= 8 ? 3 : 2);
    const int npass = (lg + maxr - 1) / maxr;
    const int base = lg / npass, extra = lg % npass;
    // Sets of pass p that are needed: walk the bound back from the last pass.
    auto kmax_of = [&](int p) {
        int bout = min(need, 1 << lg), s_hi = lg + 1, K = 0;
        for (int pp = npass - 1; pp >= p; --pp) {
            const int RL = base + (pp < extra ? 1 : 0);
            const int s_lo = s_hi - RL, ub = lg - s_lo - RL + 1;
            K = (bout + (1 << ub) - 1) >> ub;
            bout = min(K << (ub + RL), 1 << lg);
            s_hi = s_lo;
        }
        return K;
    };
    int s_lo = 1;
    for (int p = 0; p < npass; ++p) {
        const int RL = base + (p < extra ? 1 : 0);
        const int kmax = need >= (1 << lg) ? (1 << 30) : kmax_of(p);
        switch (RL) {
            case 1: tile_pass<C, 1, PRE>(buf, T, s_lo, twp, gstride, conj, kmax); break;
            case 2: tile_pass<C, 2, PRE>(buf, T, s_lo, twp, gstride, conj, kmax); break;
            case 3:
                if constexpr (C::kPT >= 8) tile_pass<C, 3, PRE>(buf, T, s_lo, twp, gstride, conj, kmax);
                break;
            default:
                if constexpr (C::kPT >= 16) tile_pass<C, 4, PRE>(buf, T, s_lo, twp, gstride, conj, kmax);
                break;
        }
        s_lo += RL;
    }
}

// Block -> (job, group block).  Workgroups are dispatched to the 8 XCDs
// round-robin by id; when possible each XCD keeps a fixed eighth of the
// group blocks for every job, so its L2 holds only that eighth of the
// per-group twiddle tables.
__device__ __forceinline__ void xcd_map(int bid, int bpj, int& job, int& gb) {
    if ((bpj & 7) == 0) {
        const int x = bid & 7, i = bid >> 3, per = bpj >> 3;
        gb = x + 8 * (i % per);
        job = i / per;
    } else {
        job = bid / bpj;
        gb = bid % bpj;
    }
}

struct FcrArgs {
    const double2* gamma;   // [nf][Nt]
    const double* omegas;   // [nf]
    const int* mirror;      // [nf]
    int nf;
    long Nt;                // Gamma column stride
    long Nin;               // Gamma rows that enter the FFT inputs (Nt, or T in the window mode)
    double dt;
    int L, a, b;
    long n;
    const double2* twStd;   // stages 1..max(a,b), standard layout
    const double2* twB;     // phase B forward stages, per k1 group
    const double2* twC;     // phase C inverse stages, per q group
    const int* job_tri;     // sorted position -> triple id (triples grouped by equal Omega)
    const unsigned char* job_mask;  // per sorted position: bit0 F needed, bit1 C needed, bit2 R needed
    long nfft_jobs;         // positions [0, nfft_jobs) need FFTs; the rest only R
    const int* ajob_start;  // [nA + 1] first sorted position of each A-job (the triples
                            // of one Omega class with the same i: identical phase-A input)
    const int* cls_start;   // [ncls + 1] first A-job of each Omega-class piece
    const int* bperm;       // phase B processing order: grid slot (per chunk) -> A-job
    long job0;              // first sorted position of the FFT chunk
    int njobs;              // triples in the FFT chunk
    long a0;                // first A-job of the FFT chunk
    int na;                 // A-jobs in the FFT chunk
    int cls0;               // first class piece touching the chunk
    long comb0;             // first sorted position of the combine batch
    long cconv;             // combine-batch capacity (conv slots)
    double2* inter1;        // [njobs][2^a][2^b]
    double2* inter2;        // [2][njobs][2^b][2^a]
    double2* spec;          // [2 nf][2^a][2^b]   spectra of g_b (even) and conj(g_b) (odd)
    double2* conv;          // [2][njobs][T]
    long T;
    // Phase mode TACO_PHASE_LIBM: host table of exp(+i Omega (t dt)), one row
    // of ph_len values per distinct Omega, and the row of each sorted
    // position; null in TACO_PHASE_CR (the kernels evaluate cexpi).
    const double2* ph_tab;
    const int* job_row;
    long ph_len;
};

// exp(+i Omega (t dt)) of a triple: its table row, or the correctly rounded value.
__device__ __forceinline__ double2 phase_plus(const double2* row, long t, double omega, double dt) {
    return row ? row[t] : cexpi(omega * (static_cast<double>(t) * dt));
}

__device__ __forceinline__ void decode_triple(long tri, int nf, int& i, int& j, int& k) {
    k = static_cast<int>(tri % nf);
    j = static_cast<int>((tri / nf) % nf);
    i = static_cast<int>(tri / (static_cast<long>(nf) * nf));
}

// Phase A of the per-bucket spectra: input g_b (even job) or conj(g_b) (odd job).
template <class Cfg>
__global__ __launch_bounds__(Cfg::kFT) void k_phase_a_spec(FcrArgs P) {
    constexpr int kFT = Cfg::kFT;
    __shared__ double2 buf[Cfg::kLds];
    const int ngroups = 1 << P.b;
    const TileLayout Tl(P.a, tile_groups<Cfg>(P.a, ngroups));
    const int G = Tl.G, gs = Tl.gs;
    int job, gb;
    xcd_map(blockIdx.x, ngroups / G, job, gb);
    const int r0b = gb * G;
    const double2* src = P.gamma + static_cast<long>(job >> 1) * P.Nt;
    const bool conj_in = (job & 1) != 0;
    const int total = G * gs;
    for (int e = threadIdx.x; e < total; e += kFT) {
        const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
        const long t = static_cast<long>(r0b + grp) + (static_cast<long>(pos) << P.b);
        double2 x = make_double2(0.0, 0.0);
        if (t < P.Nin) {
            x = src[t];
            if (conj_in) x.y = -x.y;
        }
        buf[Tl.at(grp, pos)] = x;
    }
    __syncthreads();
    tile_fft<Cfg>(buf, Tl, P.twStd, 0, false);
    double2* out = P.inter1 + static_cast<long>(job) * P.n;
    for (int e = threadIdx.x; e < total; e += kFT) {
        const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
        out[(static_cast<long>(pos) << P.b) + r0b + grp] = buf[Tl.at(grp, pos)];
    }
}

// Phase A of the triples: one workgroup per (Omega-class piece, group block).
// The phases exp(-i Omega t_k) of the tile are computed once and applied to
// the Gamma column g_i of every A-job of the piece inside the chunk; triples
// of a class that share i share the input g_i exp(-i Omega t), so one A-job
// (its first triple) computes their common half-transformed signal.
template <class Cfg>
__global__ __launch_bounds__(Cfg::kFT) void k_phase_a(FcrArgs P) {
    constexpr int kFT = Cfg::kFT;
    constexpr int kPT = Cfg::kPT;
    __shared__ double2 buf[Cfg::kLds];
    const int ngroups = 1 << P.b;
    const TileLayout Tl(P.a, tile_groups<Cfg>(P.a, ngroups));
    const int G = Tl.G, gs = Tl.gs;
    int cj, gb;
    xcd_map(blockIdx.x, ngroups / G, cj, gb);
    const int cls = P.cls0 + cj;
    const long abeg = max(static_cast<long>(P.cls_start[cls]), P.a0);
    const long aend = min(static_cast<long>(P.cls_start[cls + 1]), P.a0 + P.na);
    if (abeg >= aend) return;
    const int r0b = gb * G;
    const long pfirst = P.ajob_start[abeg];
    int i0, j0, k0;
    decode_triple(P.job_tri[pfirst], P.nf, i0, j0, k0);
    const double omega = (P.omegas[i0] + P.omegas[j0]) + P.omegas[k0];
    const double2* prow = P.ph_tab ? P.ph_tab + static_cast<long>(P.job_row[pfirst]) * P.ph_len : nullptr;
    const int total = G * gs;
    double2 ph[kPT];
#pragma unroll
    for (int q = 0; q < kPT; ++q) {
        const int e = threadIdx.x + q * kFT;
        const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
        const long t = static_cast<long>(r0b + grp) + (static_cast<long>(pos) << P.b);
        ph[q] = make_double2(0.0, 0.0);
        if (e < total && t < P.Nin) ph[q] = cconj(phase_plus(prow, t, omega, P.dt));
    }
    for (long aj = abeg; aj < aend; ++aj) {
        int i, j, k;
        decode_triple(P.job_tri[P.ajob_start[aj]], P.nf, i, j, k);
        const double2* src = P.gamma + static_cast<long>(i) * P.Nt;
#pragma unroll
        for (int q = 0; q < kPT; ++q) {
            const int e = threadIdx.x + q * kFT;
            if (e < total) {
                const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
                const long t = static_cast<long>(r0b + grp) + (static_cast<long>(pos) << P.b);
                buf[Tl.at(grp, pos)] = (t < P.Nin) ? cmul(src[t], ph[q]) : make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
        tile_fft<Cfg>(buf, Tl, P.twStd, 0, false);
        double2* out = P.inter1 + (aj - P.a0) * P.n;
        for (int e = threadIdx.x; e < total; e += kFT) {
            const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
            out[(static_cast<long>(pos) << P.b) + r0b + grp] = buf[Tl.at(grp, pos)];
        }
        __syncthreads();
    }
}

// Phase B: forward stages a+1..L; mode 1 stores the spectrum, mode 0 forms
// the F and C products and runs inverse stages 1..b on each.
template <class Cfg>
__global__ __launch_bounds__(Cfg::kFT) TACO_FFT_OCC(Cfg) void k_phase_b(FcrArgs P, int mode) {
    constexpr int kFT = Cfg::kFT;
    constexpr int kPT = Cfg::kPT;
    __shared__ double2 buf[Cfg::kLds];
    const int ngroups = 1 << P.a;
    const TileLayout Tl(P.b, tile_groups<Cfg>(P.b, ngroups));
    const int G = Tl.G, gs = Tl.gs;
    int job, gb;
    xcd_map(blockIdx.x, ngroups / G, job, gb);
    if (mode == 0) job = P.bperm[P.a0 + job] - static_cast<int>(P.a0);  // A-job within the chunk
    const int k1b = gb * G;
    const int total = G * gs;
    const double2* in = P.inter1 + static_cast<long>(job) * P.n + static_cast<long>(k1b) * gs;
    for (int e = threadIdx.x; e < total; e += kFT) buf[Tl.at(e >> Tl.lg, e & (gs - 1))] = in[e];
    __syncthreads();
    const long bstride = (1L << P.b) - 1;
    tile_fft<Cfg, true>(buf, Tl, P.twB + static_cast<long>(k1b) * bstride, bstride, false);
    if (mode == 1) {
        double2* out = P.spec + static_cast<long>(job) * P.n + static_cast<long>(k1b) * gs;
        for (int e = threadIdx.x; e < total; e += kFT) out[e] = buf[Tl.at(e >> Tl.lg, e & (gs - 1))];
        return;
    }
    double2 X[kPT];
#pragma unroll
    for (int q = 0; q < kPT; ++q) {
        const int e = threadIdx.x + q * kFT;
        if (e < total) X[q] = buf[Tl.at(e >> Tl.lg, e & (gs - 1))];
    }
    __syncthreads();
    const long pend = P.ajob_start[P.a0 + job + 1];
    if constexpr (kPT <= 4) {
        // The A-job's live (triple, path) pairs in order (path 0, F: the
        // spectrum of g_mirror(j); path 1, C: that of conj(g_j)); the next
        // pair's spectrum slice is loaded into registers while the current
        // product is transformed (tiles of 4 values per thread have the
        // registers for it).
        auto next_live = [&](long& p, int& path) {
            while (p < pend) {
                const int mask = P.job_mask[p];
                while (path < 2 && !(mask & (1 << path))) ++path;
                if (path < 2) return;
                ++p;
                path = 0;
            }
        };
        auto load_spec = [&](double2 (&Y)[kPT], long p, int path) {
            int i, j, k;
            decode_triple(P.job_tri[p], P.nf, i, j, k);
            const double2* sp = P.spec + static_cast<long>(path == 0 ? 2 * P.mirror[j] : 2 * j + 1) * P.n +
                                static_cast<long>(k1b) * gs;
#pragma unroll
            for (int q = 0; q < kPT; ++q) {
                const int e = threadIdx.x + q * kFT;
                if (e < total) Y[q] = sp[e];
            }
        };
        long p = P.ajob_start[P.a0 + job];
        int path = 0;
        next_live(p, path);
        double2 Y[kPT];
        if (p < pend) load_spec(Y, p, path);
        while (p < pend) {
            long pn = p;
            int pathn = path + 1;
            next_live(pn, pathn);
#pragma unroll
            for (int q = 0; q < kPT; ++q) {
                const int e = threadIdx.x + q * kFT;
                if (e < total) buf[Tl.at(e >> Tl.lg, e & (gs - 1))] = cmul(X[q], Y[q]);
            }
            __syncthreads();
            if (pn < pend) load_spec(Y, pn, pathn);
            tile_fft<Cfg, true>(buf, Tl, P.twStd, 0, true);
            double2* out = P.inter2 + (static_cast<long>(path) * P.njobs + (p - P.job0)) * P.n;
            for (int e = threadIdx.x; e < total; e += kFT) {
                const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
                out[(static_cast<long>(pos) << P.a) + k1b + grp] = buf[Tl.at(grp, pos)];
            }
            __syncthreads();
            p = pn;
            path = pathn;
        }
        return;
    }
    for (long p = P.ajob_start[P.a0 + job]; p < pend; ++p) {
        int i, j, k;
        decode_triple(P.job_tri[p], P.nf, i, j, k);
        const int bF = P.mirror[j];
        const int mask = P.job_mask[p];
        for (int path = 0; path < 2; ++path) {
            if (!(mask & (1 << path))) continue;
            const double2* sp = P.spec + static_cast<long>(path == 0 ? 2 * bF : 2 * j + 1) * P.n +
                                static_cast<long>(k1b) * gs;
#pragma unroll
            for (int q = 0; q < kPT; ++q) {
                const int e = threadIdx.x + q * kFT;
                if (e < total) buf[Tl.at(e >> Tl.lg, e & (gs - 1))] = cmul(X[q], sp[e]);
            }
            __syncthreads();
            tile_fft<Cfg, true>(buf, Tl, P.twStd, 0, true);
            double2* out = P.inter2 + (static_cast<long>(path) * P.njobs + (p - P.job0)) * P.n;
            for (int e = threadIdx.x; e < total; e += kFT) {
                const int grp = Tl.grp_of(e), pos = Tl.pos_of(e);
                out[(static_cast<long>(pos) << P.a) + k1b + grp] = buf[Tl.at(grp, pos)];
            }
            __syncthreads();
        }
    }
}

// Phase C: inverse stages b+1..L; writes conv[path][job][t] = dt * (y[t] / n), t < T.
// Each workgroup takes tiles blockIdx.x and blockIdx.x + gridDim.x (same XCD
// when the grid is a multiple of 8); the second tile's input is loaded into
// registers while the first is transformed.
template <class Cfg>
__global__ __launch_bounds__(Cfg::kFT) void k_phase_c(FcrArgs P, int ntiles) {
    constexpr int kFT = Cfg::kFT;
    constexpr int kPT = Cfg::kPT;
    __shared__ double2 buf[Cfg::kLds];
    const int ngroups = 1 << P.b;
    const TileLayout Tl(P.a, tile_groups<Cfg>(P.a, ngroups));
    const int G = Tl.G, gs = Tl.gs;
    const int total = G * gs;
    const long cstride = (1L << P.a) - 1;
    const double inv_n = 1.0 / static_cast<double>(P.n);
    int jpA = 0, jpB = 0, q0A = 0, q0B = 0;
    bool liveA = false, liveB = false;
    {
        const int bA = blockIdx.x, bB = blockIdx.x + gridDim.x;
        int gb;
        if (bA < ntiles) {
            xcd_map(bA, ngroups / G, jpA, gb);  // jp = path * njobs + job
            q0A = gb * G;
            liveA = q0A < P.T && (P.job_mask[P.job0 + jpA % P.njobs] & (1 << (jpA / P.njobs)));
        }
        if (bB < ntiles) {
            xcd_map(bB, ngroups / G, jpB, gb);
            q0B = gb * G;
            liveB = q0B < P.T && (P.job_mask[P.job0 + jpB % P.njobs] & (1 << (jpB / P.njobs)));
        }
    }
    double2 vA[kPT], vB[kPT];
#pragma unroll
    for (int q = 0; q < kPT; ++q) {
        const int e = threadIdx.x + q * kFT;
        vA[q] = (liveA && e < total) ? P.inter2[static_cast<long>(jpA) * P.n + static_cast<long>(q0A) * gs + e]
                                     : make_double2(0.0, 0.0);
        vB[q] = (liveB && e < total) ? P.inter2[static_cast<long>(jpB) * P.n + static_cast<long>(q0B) * gs + e]
                                     : make_double2(0.0, 0.0);
    }
#define TACO_PHASE_C_TILE(JP, Q0, V)                                                                         \
    {                                                                                                         \
        _Pragma("unroll") for (int q = 0; q < kPT; ++q) {                                                     \
            const int e = threadIdx.x + q * kFT;                                                              \
            if (e < total) buf[Tl.at(e >> Tl.lg, e & (gs - 1))] = V[q];                                       \
        }                                                                                                     \
        __syncthreads();                                                                                      \
        /* only rows v2 with (Q0) + v2 * 2^b < T are kept */                                                 \
        const int vrows = static_cast<int>(min(static_cast<long>(gs), (P.T - (Q0) + (1L << P.b) - 1) >> P.b));  \
        tile_fft<Cfg>(buf, Tl, P.twC + static_cast<long>(Q0) * cstride, cstride, true, vrows);                \
        const int path = (JP) / P.njobs, job = (JP) % P.njobs;                                                \
        double2* out = P.conv + (path * P.cconv + (P.job0 + job - P.comb0)) * P.T;                            \
        const int etot = min(total, G * vrows);                                                               \
        for (int e = threadIdx.x; e < etot; e += kFT) {                                                       \
            const int grp = Tl.grp_of(e), v2 = Tl.pos_of(e);                                                  \
            const long tt = static_cast<long>((Q0) + grp) + (static_cast<long>(v2) << P.b);                   \
            if (tt < P.T) {                                                                                   \
                const double2 y = cscale(buf[Tl.at(grp, v2)], inv_n);                                         \
                out[tt] = make_double2(P.dt * y.x, P.dt * y.y);                                               \
            }                                                                                                 \
        }                                                                                                     \
        __syncthreads();                                                                                      \
    }
    if (liveA) TACO_PHASE_C_TILE(jpA, q0A, vA)
    if (liveB) TACO_PHASE_C_TILE(jpB, q0B, vB)
#undef TACO_PHASE_C_TILE
}

// Inclusive block scan (256 threads = 4 waves of 64) of one double.
__device__ __forceinline__ double block_scan(double v, double* wsum, double& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const double nb = __shfl_up(v, off, 64);
        if (lane >= off) v += nb;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    double pre = 0.0;
    for (int w = 0; w < wid; ++w) pre += wsum[w];
    total = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    __syncthreads();
    return pre + v;
}

// The block scan above for Q values at once: the same wave scans, prefix
// sums and totals (so bit-identical per value), two barriers in all.
template <int Q>
__device__ __forceinline__ void block_scan_n(double (&v)[Q], double (*wsum)[4], double (&total)[Q]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < Q; ++q)
        for (int off = 1; off < 64; off <<= 1) {
            const double nb = __shfl_up(v[q], off, 64);
            if (lane >= off) v[q] += nb;
        }
    if (lane == 63)
#pragma unroll
        for (int q = 0; q < Q; ++q) wsum[q][wid] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double pre = 0.0;
        for (int w = 0; w < wid; ++w) pre += wsum[q][w];
        total[q] = (wsum[q][0] + wsum[q][1]) + (wsum[q][2] + wsum[q][3]);
        v[q] = pre + v[q];
    }
    __syncthreads();
}

struct CombineArgs {
    int out_mode;        // 0: time-major slices at t2pos; 1: series-major [tri][T]
    const int* t2pos;    // [T], -1 when t is not requested (mode 0)
    long npos;           // number of kept times (mode 0)
    long nf3;
    double2* Fo;
    double2* Co;
    double2* Ro;
    long tri0;           // mode 1: series of triple tri go to row tri - tri0
};

// F/C/R assembly for t < T (tcl4_kernels.cpp:105-175): prefix integrals by
// block scans carried across 256-sample blocks, combined with the FFT
// convolution terms of phase C.  combine_block handles one block of one
// triple: the eight scanned quantities (F: g2m pp; C: conj(g2) pp; R: g2 pm
// and g1 g2 pm), their block totals, and (WRITE) the outputs from the
// carried prefix.
struct CombineTriple {
    long cpos, tri;
    int mask;
    bool has_conv;
    const double2 *g1, *g2, *g2m, *convF, *convC, *prow;
    double omega;
};

__device__ __forceinline__ CombineTriple combine_triple(const FcrArgs& P, int slot) {
    CombineTriple c;
    c.cpos = P.comb0 + slot;
    c.tri = P.job_tri[c.cpos];
    c.mask = P.job_mask[c.cpos];
    c.has_conv = c.cpos < P.nfft_jobs;
    int i, j, k;
    decode_triple(c.tri, P.nf, i, j, k);
    c.omega = (P.omegas[i] + P.omegas[j]) + P.omegas[k];
    c.prow = P.ph_tab ? P.ph_tab + static_cast<long>(P.job_row[c.cpos]) * P.ph_len : nullptr;
    c.g1 = P.gamma + static_cast<long>(i) * P.Nt;
    c.g2 = P.gamma + static_cast<long>(j) * P.Nt;
    c.g2m = P.gamma + static_cast<long>(P.mirror[j]) * P.Nt;
    c.convF = P.conv + static_cast<long>(slot) * P.T;
    c.convC = P.conv + (P.cconv + slot) * P.T;
    return c;
}

template <bool WRITE>
__device__ __forceinline__ void combine_block(const FcrArgs& P, const CombineArgs& O, const CombineTriple& c, long t0,
                                              const double (&carry)[8], double (&tot)[8], double (*wsum)[4]) {
    const long t = t0 + threadIdx.x;
    const bool valid = t < P.T;
    double2 a1 = make_double2(0, 0), a2 = a1, a3 = a1, a4 = a1, pm = a1, gg1 = a1;
    if (valid) {
        const double2 pp = phase_plus(c.prow, t, c.omega, P.dt);
        pm = cconj(pp);
        gg1 = c.g1[t];
        const double2 gg2 = c.g2[t];
        a1 = cmul(c.g2m[t], pp);        // F: g2m * phase_plus
        a2 = cmul(cconj(gg2), pp);      // C: conj(g2) * phase_plus
        a3 = cmul(gg2, pm);             // R: g2 * phase_minus
        a4 = cmul(cmul(gg1, gg2), pm);  // R: (g1 g2) * phase_minus
    }
    double v[8] = {a1.x, a1.y, a2.x, a2.y, a3.x, a3.y, a4.x, a4.y};
    block_scan_n<8>(v, wsum, tot);
    if constexpr (!WRITE) return;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = v[q] + carry[q];
    if (!valid) return;
    long dst = -1;
    if (O.out_mode == 0) {
        const int pos = O.t2pos[t];
        if (pos < 0) return;
        dst = static_cast<long>(pos) * O.nf3 + c.tri;
    } else {
        dst = (c.tri - O.tri0) * P.T + t;
    }
    const double dt = P.dt;
    const double2 pF = make_double2(v[0] * dt, v[1] * dt);
    const double2 pC = make_double2(v[2] * dt, v[3] * dt);
    const double2 pR = make_double2(v[4] * dt, v[5] * dt);
    const double2 pP = make_double2(v[6] * dt, v[7] * dt);
    const double2 g1p = cmul(gg1, pm);
    if (c.has_conv && (c.mask & 1)) O.Fo[dst] = csub(cmul(g1p, pF), c.convF[t]);
    if (c.has_conv && (c.mask & 2)) O.Co[dst] = csub(cmul(g1p, pC), c.convC[t]);
    if (c.mask & 4) O.Ro[dst] = csub(cmul(gg1, pR), pP);
}

// Direct method (tcl4_kernels.cpp:188-281), the reference's O(T^2)
// cross-check of the convolution route: one thread per (triple, t) forms
// F(t), C(t) and R(t) from their defining sums, each sum taken in the
// reference's order (m = 0, 1, ..., t) with the reference's grouping of
// products, so the values are those of the sequential code.
struct DirectArgs {
    const double2* gamma;  // [nf][Nt]
    long Nt;
    int nf;
    const double* omegas;
    const int* mirror;
    double dt;
    long T;                // times t < T
    const int* tri;        // [ntri] triples, ascending
    const unsigned char* mask;
    const double2* ph_tab;  // TACO_PHASE_LIBM rows (null: correctly rounded on the device)
    const int* row;
    long ph_len;
};

__global__ __launch_bounds__(kBlock) void k_fcr_direct(DirectArgs D, CombineArgs O) {
    const long q = blockIdx.x;
    const long t = static_cast<long>(blockIdx.y) * kBlock + threadIdx.x;
    if (t >= D.T) return;
    long dst = -1;
    const int tri = D.tri[q];
    if (O.out_mode == 0) {
        const int pos = O.t2pos[t];
        if (pos < 0) return;
        dst = static_cast<long>(pos) * O.nf3 + tri;
    } else {
        dst = (tri - O.tri0) * D.T + t;
    }
    const int mask = D.mask[q];
    int i, j, k;
    decode_triple(tri, D.nf, i, j, k);
    const double omega = (D.omegas[i] + D.omegas[j]) + D.omegas[k];
    const double2* prow = D.ph_tab ? D.ph_tab + static_cast<long>(D.row[q]) * D.ph_len : nullptr;
    const double2* g1 = D.gamma + static_cast<long>(i) * D.Nt;
    const double2* g2 = D.gamma + static_cast<long>(j) * D.Nt;
    const double2* g2m = D.gamma + static_cast<long>(D.mirror[j]) * D.Nt;
    const double dt = D.dt;
    const double2 pm_t = cconj(phase_plus(prow, t, omega, dt));
    const double2 z = make_double2(0.0, 0.0);
    if (mask & 3) {
        // F: g2 at mirror(j); C: conj(g2) at j (tcl4.cpp:128-150)
        double2 preF = z, preC = z, secF = z, secC = z;
        for (long m = 0; m <= t; ++m) {
            const double2 pp = phase_plus(prow, m, omega, dt);
            const double2 pmk = cconj(phase_plus(prow, t - m, omega, dt));
            const double2 a = g1[t - m];
            if (mask & 1) {
                preF = cadd(preF, cscale(cmul(g2m[m], pp), dt));
                secF = cadd(secF, cmul(cmul(a, g2m[m]), pmk));
            }
            if (mask & 2) {
                const double2 b = cconj(g2[m]);
                preC = cadd(preC, cscale(cmul(b, pp), dt));
                secC = cadd(secC, cmul(cmul(a, b), pmk));
            }
        }
        if (mask & 1) O.Fo[dst] = csub(cmul(cmul(g1[t], preF), pm_t), cscale(secF, dt));
        if (mask & 2) O.Co[dst] = csub(cmul(cmul(g1[t], preC), pm_t), cscale(secC, dt));
    }
    if (mask & 4) {
        double2 p1 = z, p2 = z;
        for (long m = 0; m <= t; ++m) {
            const double2 pm = cconj(phase_plus(prow, m, omega, dt));
            p1 = cadd(p1, cscale(cmul(g2[m], pm), dt));
            p2 = cadd(p2, cscale(cmul(cmul(g1[m], g2[m]), pm), dt));
        }
        O.Ro[dst] = csub(cmul(g1[t], p1), p2);
    }
}

// One workgroup per triple walks the blocks t0 = 0, 256, ... with the carry.
__global__ __launch_bounds__(kBlock) void k_combine(FcrArgs P, CombineArgs O) {
    __shared__ double wsum[8][4];
    const CombineTriple c = combine_triple(P, static_cast<int>(blockIdx.x));
    double carry[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long t0 = 0; t0 < P.T; t0 += kBlock) {
        double tot[8];
        combine_block<true>(P, O, c, t0, carry, tot, wsum);
#pragma unroll
        for (int q = 0; q < 8; ++q) carry[q] += tot[q];
    }
}

// The same in three launches when the triples are too few to fill the chip:
// block totals (a workgroup per (triple, block)), the carries into every
// block by the sequential sum above (a thread per (triple, quantity)), then
// every block with its carry.  Each value is formed by the same operations
// in the same order, so the outputs equal k_combine's bit for bit.
__global__ __launch_bounds__(kBlock) void k_combine_tot(FcrArgs P, CombineArgs O, double* tot_out, int nblk) {
    __shared__ double wsum[8][4];
    const CombineTriple c = combine_triple(P, static_cast<int>(blockIdx.x));
    const double carry[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double tot[8];
    combine_block<false>(P, O, c, static_cast<long>(blockIdx.y) * kBlock, carry, tot, wsum);
    if (threadIdx.x < 8) tot_out[(static_cast<long>(blockIdx.x) * nblk + blockIdx.y) * 8 + threadIdx.x] = tot[threadIdx.x];
}

__global__ void k_combine_carry(double* tot_carry, long cb, int nblk) {
    const long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;  // (slot, quantity)
    if (e >= cb * 8) return;
    const long slot = e / 8, q = e % 8;
    double* p = tot_carry + slot * nblk * 8 + q;
    double carry = 0.0;
    for (int b = 0; b < nblk; ++b) {
        const double tb = p[static_cast<long>(b) * 8];
        p[static_cast<long>(b) * 8] = carry;
        carry += tb;
    }
}

__global__ __launch_bounds__(kBlock) void k_combine_blocks(FcrArgs P, CombineArgs O, const double* carry_in, int nblk) {
    __shared__ double wsum[8][4];
    const CombineTriple c = combine_triple(P, static_cast<int>(blockIdx.x));
    double carry[8];
    const double* ci = carry_in + (static_cast<long>(blockIdx.x) * nblk + blockIdx.y) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) carry[q] = ci[q];
    double tot[8];
    combine_block<true>(P, O, c, static_cast<long>(blockIdx.y) * kBlock, carry, tot, wsum);
}

__global__ void k_cr_sincos(const double* x, long n, double* s, double* c) {
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (i < n) taco_cr::sincos_cr(x[i], s + i, c + i, c_crtab);
}

__global__ void k_cr_exp(const double* x, long n, double* out) {
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (i < n) out[i] = taco_cr::exp_cr(x[i], c_exptab);
}

// ---------------------------------------------------------------------------
// Bath correlation function and trapz Gamma on the device: the twins of the
// host producers (bcf_fft_fun, correlation_fft.hpp:154-214, and
// compute_trapz_prefix_multi_matrix, gamma.hpp:166-213), bit-identical to
// them.  Their two recurrences (the KMS factor r^k by repeated products and
// the trapz phase / sum chains) are sequential in the reference's arithmetic,
// so one lane runs each chain; everything else is parallel.
// ---------------------------------------------------------------------------
struct BcfArgs {
    int kind;
    double p0, p1, beta;   // beta = +inf: zero temperature
    double domega, omegaNy, r, eNy;
    long nfft;
    const double* kms;     // [nfft / 2] KMS factor used at k (finite beta)
    double2* S;            // [nfft] sampled spectrum
    // TACO_PHASE_LIBM, Ohmic J: the host's std::exp(-w_k / p1) at w_k =
    // domega k (k <= nfft/2 - 1) and at omegaNy; null otherwise.
    const double* efac;
    double efac_ny;
};

// J(w); ef < 0 asks for the correctly rounded Ohmic factor, else ef is it.
__device__ __forceinline__ double spectral_J_dev(int kind, double p0, double p1, double w, double ef) {
    if (!(w > 0.0)) return 0.0;
    if (kind == 0) return p0 * w * (ef >= 0.0 ? ef : taco_cr::exp_cr(-w / p1, c_exptab));
    return 2.0 * p0 * p1 * w / (w * w + p1 * p1);
}

// kms[k] = r^k for k = 1..posCount by the reference's recurrence KMS *= r.
// One lane; the chain is unrolled 32 deep (loop overhead, not the multiply's
// latency, bounds a shallower loop: 9.3 -> 6.2 ns per factor).
__global__ void k_bcf_kms(double r, long pos_count, double* kms) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    constexpr int kU = 32;
    double v = r;
    long k = 1;
    for (; k + kU <= pos_count + 1; k += kU) {  // the same chain, stored kU at a time
        double t[kU];
#pragma unroll
        for (int j = 0; j < kU; ++j) {
            t[j] = v;
            v *= r;
        }
#pragma unroll
        for (int j = 0; j < kU; ++j) kms[k + j] = t[j];
    }
    for (; k <= pos_count; ++k) {
        kms[k] = v;
        v *= r;
    }
}

// S[k], k < nfft: the KMS spectrum of bcf_fft_fun, one entry per thread.
__global__ __launch_bounds__(kBlock) void k_bcf_spectrum(BcfArgs B) {
    const long k = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x;
    if (k >= B.nfft) return;
    const long half = B.nfft / 2, pos_count = half - 1;
    constexpr double kPiD = 3.14159265358979323846;
    double v = 0.0;
    const bool zero_t = isinf(B.beta);
    if (k == half) {
        const double J = spectral_J_dev(B.kind, B.p0, B.p1, B.omegaNy, B.efac ? B.efac_ny : -1.0);
        v = zero_t ? 0.25 * kPiD * J : 0.25 * kPiD * J * (1.0 + 2.0 / (B.eNy - 1.0));
    } else if (k == 0) {
        if (!zero_t) {
            const double ed = B.r;
            v = 0.5 * (kPiD / 2.0) * spectral_J_dev(B.kind, B.p0, B.p1, B.domega, B.efac ? B.efac[1] : -1.0) *
                (1.0 + ed) / (1.0 - ed);
        }
    } else if (k <= pos_count) {
        const double w = B.domega * static_cast<double>(k);
        const double ef = B.efac ? B.efac[k] : -1.0;
        v = zero_t ? (kPiD / 2.0) * spectral_J_dev(B.kind, B.p0, B.p1, w, ef)
                   : (kPiD / 2.0) * spectral_J_dev(B.kind, B.p0, B.p1, w, ef) / (1.0 - B.kms[k]);
    } else if (!zero_t) {
        const long kk = B.nfft - k;  // 1..posCount
        const double w = B.domega * static_cast<double>(kk);
        const double Sp = (kPiD / 2.0) * spectral_J_dev(B.kind, B.p0, B.p1, w, B.efac ? B.efac[kk] : -1.0) /
                          (1.0 - B.kms[kk]);
        v = Sp * B.kms[kk];
    }
    B.S[k] = make_double2(v, 0.0);
}

// The reference's radix-2 FFT (bit reversal, then each stage's butterflies
// u +- w x with the recurrence twiddles of the host table) on one thread, for
// the lengths 2 and 4 below the three-pass kernels' range (bath grids with
// N <= 1).  tw: stage s at offset 2^(s-1) - 1.
__global__ void k_fft_tiny(double2* a, int L, const double2* tw) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int n = 1 << L;
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            const double2 t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int half = 1; half < n; half <<= 1) {
        const double2* w = tw + (half - 1);
        for (int i = 0; i < n; i += 2 * half)
            for (int k = 0; k < half; ++k) {
                const double2 u = a[i + k];
                const double2 v = cmul(w[k], a[i + k + half]);
                a[i + k] = cadd(u, v);
                a[i + k + half] = csub(u, v);
            }
    }
}

// C[n] = scale * X[n], n < nc, from the phase-B spectrum layout
// spec[k1 * 2^b + p] = X[k1 + p * 2^a].
__global__ __launch_bounds__(kBlock) void k_bcf_gather(const double2* spec, int a, int b, long nc, double scale,
                                                       double2* C) {
    const long n = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x;
    if (n >= nc) return;
    const long src = ((n & ((1L << a) - 1)) << b) + (n >> a);
    const double2 x = spec[src];
    C[n] = make_double2(x.x * scale, x.y * scale);
}

// Gamma[j][k] = trapz prefix of exp(i omega_j tau) C(tau), one workgroup per
// bucket, with the reference's chains (phi_k = phi_{k-1} step; acc_k =
// acc_{k-1} + dt/2 (phi_{k-1} C_{k-1} + phi_k C_k)) split so that only the
// two recurrences are sequential.  Per chunk of kTzChunk times, in a
// pipeline: lane 0 of wave 0 runs the phase chain of chunk c, lane 0 of wave
// 1 the sum chain of chunk c - 1, waves 2-3 load C of chunk c and write Gamma
// of chunk c - 2; then every thread forms chunk c's terms, each exactly the
// reference's expression.  Over full chunks the chains run unrolled 64 (phase)
// and 32 (sum) deep: a shallower loop's overhead rivals the arithmetic (5
// buckets x 65536 times: 2.12 -> 1.99 ms).
constexpr int kTzChunk = 1024;
static_assert(kTzChunk % (kBlock - 128) == 0, "loader waves cover a chunk exactly");
static_assert(kTzChunk % 64 == 0, "unrolled chains cover a full chunk exactly");
__global__ __launch_bounds__(kBlock) void k_gamma_trapz(const double2* C, long nc, double dt, const double* omegas,
                                                        const double2* steps, double2* G) {
    __shared__ double2 cin[2][kTzChunk];   // C of chunk c (c % 2)
    __shared__ double2 ph[2][kTzChunk];    // phi_k of chunk c
    __shared__ double2 uu[2][kTzChunk];    // terms of chunk c
    __shared__ double2 gout[2][kTzChunk];  // Gamma of chunk c
    __shared__ double2 last_phi[2], last_c[2];
    double2* col = G + static_cast<long>(blockIdx.x) * nc;
    const double2 half_dt = make_double2(dt / 2.0, 0.0);
    double2 step = make_double2(1.0, 0.0), phi = make_double2(1.0, 0.0), acc = make_double2(0.0, 0.0);
    // the phase step exp(i omega_b dt): the host's value (TACO_PHASE_LIBM) or correctly rounded
    if (threadIdx.x == 0) step = steps ? steps[blockIdx.x] : cexpi(omegas[blockIdx.x] * dt);
    const long nch = (nc + kTzChunk - 1) / kTzChunk;
    for (long c = 0; c <= nch + 1; ++c) {
        const int cb = static_cast<int>(c & 1), pb = cb ^ 1;
        if (threadIdx.x == 0) {
            if (c < nch) {  // phase chain of chunk c
                const long k0 = c * kTzChunk;
                const int e_end = static_cast<int>(min(static_cast<long>(kTzChunk), nc - k0));
                int e = 0;
                if (k0 == 0) {
                    ph[cb][0] = phi;  // phi_0 = 1
                    e = 1;
                }
                if (e_end == kTzChunk) {  // full chunk: to a multiple of 64, then 64 deep
                    for (; e % 64 != 0; ++e) {
                        phi = cmul(phi, step);
                        ph[cb][e] = phi;
                    }
                    for (; e < kTzChunk; e += 64) {
#pragma unroll
                        for (int j = 0; j < 64; ++j) {
                            phi = cmul(phi, step);
                            ph[cb][e + j] = phi;
                        }
                    }
                }
                for (; e < e_end; ++e) {
                    phi = cmul(phi, step);
                    ph[cb][e] = phi;
                }
            }
        } else if (threadIdx.x == 64) {
            if (c >= 1 && c - 1 < nch) {  // sum chain of chunk c - 1
                const long k0 = (c - 1) * kTzChunk;
                const int e_end = static_cast<int>(min(static_cast<long>(kTzChunk), nc - k0));
                int e = 0;
                if (k0 == 0) {
                    gout[pb][0] = make_double2(0.0, 0.0);
                    e = 1;
                }
                if (e_end == kTzChunk) {  // full chunk: to a multiple of 32, then 32 deep
                    for (; e % 32 != 0; ++e) {
                        acc = cadd(acc, uu[pb][e]);
                        gout[pb][e] = acc;
                    }
                    for (; e < kTzChunk; e += 32) {
                        double2 u[32];
#pragma unroll
                        for (int j = 0; j < 32; ++j) u[j] = uu[pb][e + j];
#pragma unroll
                        for (int j = 0; j < 32; ++j) {
                            acc = cadd(acc, u[j]);
                            gout[pb][e + j] = acc;
                        }
                    }
                }
                for (; e < e_end; ++e) {
                    acc = cadd(acc, uu[pb][e]);
                    gout[pb][e] = acc;
                }
            }
        } else if (threadIdx.x >= 128) {
            // kLd values per loader thread, all loads issued before the LDS writes
            constexpr int kLd = kTzChunk / (kBlock - 128);
            const int e0 = threadIdx.x - 128;
            if (c < nch) {
                double2 v[kLd];
#pragma unroll
                for (int q = 0; q < kLd; ++q) {
                    const long k = c * kTzChunk + e0 + q * (kBlock - 128);
                    v[q] = k < nc ? C[k] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int q = 0; q < kLd; ++q) cin[cb][e0 + q * (kBlock - 128)] = v[q];
            }
            if (c >= 2) {
                double2 v[kLd];
#pragma unroll
                for (int q = 0; q < kLd; ++q) v[q] = gout[cb][e0 + q * (kBlock - 128)];
#pragma unroll
                for (int q = 0; q < kLd; ++q) {
                    const long k = (c - 2) * kTzChunk + e0 + q * (kBlock - 128);
                    if (k < nc) col[k] = v[q];
                }
            }
        }
        __syncthreads();
        if (c < nch) {  // terms of chunk c
            const long k0 = c * kTzChunk;
            const int e_end = static_cast<int>(min(static_cast<long>(kTzChunk), nc - k0));
            for (int e = threadIdx.x; e < e_end; e += kBlock) {
                if (k0 + e == 0) continue;
                const double2 pm = e > 0 ? ph[cb][e - 1] : last_phi[pb];
                const double2 cm = e > 0 ? cin[cb][e - 1] : last_c[pb];
                uu[cb][e] = cmul(half_dt, cadd(cmul(pm, cm), cmul(ph[cb][e], cin[cb][e])));
            }
            if (threadIdx.x == 0 && e_end == kTzChunk) {
                last_phi[cb] = ph[cb][kTzChunk - 1];
                last_c[cb] = cin[cb][kTzChunk - 1];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Fused MIKX + GW + L4 for one time index per workgroup.
// ---------------------------------------------------------------------------
struct L4Args {
    const double2* Fs;   // [npos][nf3]
    const double2* Cs;
    const double2* Rs;
    const int* p2f;      // N x N column-major
    const double2* ops;  // num_ops x N x N
    int N, nf, num_ops;
    long nf3;
    double2* L4;         // [npos][N^4]
    double2* GW;         // optional [npos][N^4]
};

__global__ __launch_bounds__(kBlock) void k_l4(L4Args A) {
    extern __shared__ double2 sm[];
    const int N = A.N, N2 = N * N, N4 = N2 * N2;
    double2* Tm = sm;                       // N^4
    double2* opsS = sm + N4;                // num_ops * N^2
    int* f = reinterpret_cast<int*>(opsS + A.num_ops * N2);  // N^2
    const long pos = blockIdx.x;
    for (int e = threadIdx.x; e < A.num_ops * N2; e += kBlock) opsS[e] = A.ops[e];
    for (int e = threadIdx.x; e < N2; e += kBlock) f[e] = A.p2f[e];
    __syncthreads();
    const double2* Fp = A.Fs + pos * A.nf3;
    const double2* Cp = A.Cs + pos * A.nf3;
    const double2* Rp = A.Rs + pos * A.nf3;
    const long nf = A.nf;
    auto id3 = [nf](int x, int y, int z) { return (static_cast<long>(x) * nf + y) * nf + z; };
    auto F2 = [f, N](int r, int c) { return f[r + c * N]; };
    for (int idx = threadIdx.x; idx < N4; idx += kBlock) {
        const int row_u = idx % N2, col_u = idx / N2;
        const int n = row_u % N, i = row_u / N, m = col_u % N, j = col_u / N;
        double2 res = make_double2(0.0, 0.0);
        for (int iA = 0; iA < A.num_ops; ++iA) {
            const double2* Ap = opsS + iA * N2;
            for (int iB = 0; iB < A.num_ops; ++iB) {
                const double2* Bp = opsS + iB * N2;
#define AA(r, c) Ap[(r) + (c) * N]
#define BB(r, c) Bp[(r) + (c) * N]
                for (int a = 0; a < N; ++a) {
                    for (int b = 0; b < N; ++b) {
                        // M(b,i;n,a) = F[f(b,a)][f(b,i)][f(n,b)] - R[f(b,a)][f(n,a)][f(a,i)]
                        const double2 Mv = csub(Fp[id3(F2(b, a), F2(b, i), F2(n, b))],
                                                Rp[id3(F2(b, a), F2(n, a), F2(a, i))]);
                        // I(a,n;i,b) = F[f(a,n)][f(b,i)][f(n,b)]
                        const double2 Iv = Fp[id3(F2(a, n), F2(b, i), F2(n, b))];
                        // K(i,b;n,a) = R[f(i,b)][f(n,a)][f(a,i)]
                        const double2 Kv = Rp[id3(F2(i, b), F2(n, a), F2(a, i))];
                        const long x1 = id3(F2(a, n), F2(j, b), F2(n, i));
                        const long x2 = id3(F2(i, a), F2(j, b), F2(n, i));
                        const long x3 = id3(F2(b, a), F2(j, m), F2(a, i));
                        const long x4 = id3(F2(i, b), F2(j, m), F2(a, i));
                        const double2 X1 = cadd(Cp[x1], Rp[x1]);
                        const double2 X2 = cadd(Cp[x2], Rp[x2]);
                        const double2 X3 = cadd(Cp[x3], Rp[x3]);
                        const double2 X4 = cadd(Cp[x4], Rp[x4]);
                        const double2 t1 = cmul(cmul(cmul(cmul(BB(n, a), AA(a, b)), BB(b, i)), AA(j, m)), Mv);
                        const double2 t2 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), BB(b, i)), AA(j, m)), Iv);
                        const double2 t3 = cmul(cmul(cmul(cmul(BB(n, a), BB(a, b)), AA(b, i)), AA(j, m)), Kv);
                        const double2 t4 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, i)), BB(j, b)), AA(b, m)), X1);
                        const double2 t5 = cmul(cmul(cmul(cmul(BB(n, a), AA(a, i)), BB(j, b)), AA(b, m)), X2);
                        const double2 t6 = cmul(cmul(cmul(cmul(AA(n, a), AA(a, b)), BB(b, i)), BB(j, m)), X3);
                        const double2 t7 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), AA(b, i)), BB(j, m)), X4);
                        double2 s = csub(t1, t2);
                        s = cadd(s, t3);
                        s = cadd(s, t4);
                        s = csub(s, t5);
                        s = csub(s, t6);
                        s = cadd(s, t7);
                        res = csub(res, s);
                    }
                }
                if (j == m) {
                    for (int a = 0; a < N; ++a)
                        for (int b = 0; b < N; ++b)
                            for (int c = 0; c < N; ++c) {
                                // M(c,i;a,b) = F[f(c,b)][f(c,i)][f(a,c)] - R[f(c,b)][f(a,b)][f(b,i)]
                                const double2 Mv = csub(Fp[id3(F2(c, b), F2(c, i), F2(a, c))],
                                                        Rp[id3(F2(c, b), F2(a, b), F2(b, i))]);
                                // K(i,c;a,b) = R[f(i,c)][f(a,b)][f(b,i)]
                                const double2 Kv = Rp[id3(F2(i, c), F2(a, b), F2(b, i))];
                                // I(b,a;i,c) = F[f(b,a)][f(c,i)][f(a,c)]
                                const double2 Iv = Fp[id3(F2(b, a), F2(c, i), F2(a, c))];
                                const double2 add1 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), AA(b, c)), BB(c, i)), Mv);
                                const double2 add2 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), BB(b, c)), AA(c, i)), Kv);
                                const double2 nA = make_double2(-AA(n, a).x, -AA(n, a).y);
                                const double2 add3 = cmul(cmul(cmul(cmul(nA, AA(a, b)), BB(b, c)), BB(c, i)), Iv);
                                res = cadd(res, cadd(cadd(add1, add2), add3));
                            }
                }
#undef AA
#undef BB
            }
        }
        Tm[row_u + col_u * N2] = res;
    }
    __syncthreads();
    double2* L = A.L4 + pos * N4;
    for (int idx = threadIdx.x; idx < N4; idx += kBlock) {
        const int r = idx % N2, c = idx / N2;
        const double2 g = cadd(Tm[r + c * N2], cconj(Tm[c + r * N2]));
        if (A.GW) A.GW[pos * N4 + idx] = g;
        const int n = r % N, i = r / N, m = c % N, j = c / N;
        L[(n + m * N) + (i + j * N) * N2] = g;
    }
}

// ---------------------------------------------------------------------------
// Device-level helpers of the reference's split path, for callers that keep
// it: the explicit M/I/K (N^4) and X (N^6) tensors at one time from the nf^3
// slices (build_mikx_device, tcl4_mikx_cuda.hpp:44-46; the gather of
// tcl4_mikx.cpp:101-132 with the code's M1 order) and GW from them
// (assemble_liouvillian_cuda_device[_raw], tcl4_assemble_cuda.hpp:24-44).
// The fused kernels above never materialise X; these exist for drop-in use.
// ---------------------------------------------------------------------------
// One thread per X element (j,k,p,q,r,s), flat6 = j + N(k + N(p + N(q + N(r + N s))));
// the (r,s) = (0,0) thread also writes M, I, K at (row j + N k, col p + N q).
__global__ __launch_bounds__(kBlock) void k_mikx(const double2* Fs, const double2* Cs, const double2* Rs,
                                                 const int* p2f, int N, int nf, double2* M, double2* I, double2* K,
                                                 double2* X) {
    extern __shared__ int fmap[];
    const int N2 = N * N;
    int bad = 0;
    for (int e = threadIdx.x; e < N2; e += kBlock) {
        fmap[e] = p2f[e];
        bad |= (fmap[e] < 0 || fmap[e] >= nf);
    }
    const bool invalid = __syncthreads_or(bad);
    const long N4 = static_cast<long>(N2) * N2;
    const long x = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x;
    if (x >= N4 * N2) return;
    const double nan = __builtin_nan("");
    const long outer = x % N4;      // j + N k + N^2 (p + N q)
    const int rs = static_cast<int>(x / N4);
    const int j = static_cast<int>(outer % N), k = static_cast<int>((outer / N) % N);
    const int p = static_cast<int>((outer / N2) % N), q = static_cast<int>(outer / (static_cast<long>(N2) * N));
    const int r = rs % N, s = rs / N;
    if (invalid) {  // a bucket index outside [0, nf): NaN instead of reads outside the slices
        X[x] = make_double2(nan, nan);
        if (rs == 0) M[outer] = I[outer] = K[outer] = make_double2(nan, nan);
        return;
    }
    auto f = [&](int a, int b) { return static_cast<long>(fmap[a + b * N]); };
    auto id3 = [nf](long a, long b, long c) { return (a * nf + b) * nf + c; };
    const long e3 = id3(f(j, k), f(p, q), f(r, s));
    X[x] = cadd(Cs[e3], Rs[e3]);
    if (rs == 0) {
        M[outer] = csub(Fs[id3(f(j, q), f(j, k), f(p, j))], Rs[id3(f(j, q), f(p, q), f(q, k))]);
        I[outer] = Fs[id3(f(j, k), f(q, p), f(k, q))];
        K[outer] = Rs[id3(f(j, k), f(p, q), f(q, j))];
    }
}

// T(n,i;m,j) of the NAKZWAN sums (tcl4_assemble.cpp:29-178) from explicit
// M/I/K/X, one thread per element, the terms in the reference's order.
__global__ __launch_bounds__(kBlock) void k_gw_raw(const double2* M, const double2* I, const double2* K,
                                                   const double2* X, const double2* ops, int N, int num_ops,
                                                   double2* GW) {
    const int N2 = N * N;
    const long idx = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x;
    if (idx >= static_cast<long>(N2) * N2) return;
    const int row_u = static_cast<int>(idx % N2), col_u = static_cast<int>(idx / N2);
    const int n = row_u % N, i = row_u / N, m = col_u % N, j = col_u / N;
    auto mat = [N2](const double2* Mx, int r, int c) { return Mx[r + static_cast<long>(c) * N2]; };
    auto l2 = [N](int r, int c) { return r + c * N; };
    auto l6 = [N](int a, int b, int c, int d, int e, int f) {
        return a + static_cast<long>(N) * (b + static_cast<long>(N) * (c + static_cast<long>(N) *
               (d + static_cast<long>(N) * (e + static_cast<long>(N) * f))));
    };
    double2 res = make_double2(0.0, 0.0);
    for (int iA = 0; iA < num_ops; ++iA) {
        const double2* Ap = ops + static_cast<long>(iA) * N2;
        for (int iB = 0; iB < num_ops; ++iB) {
            const double2* Bp = ops + static_cast<long>(iB) * N2;
#define AA(r, c) Ap[(r) + (c) * N]
#define BB(r, c) Bp[(r) + (c) * N]
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) {
                    const double2 t1 = cmul(cmul(cmul(cmul(BB(n, a), AA(a, b)), BB(b, i)), AA(j, m)),
                                            mat(M, l2(b, i), l2(n, a)));
                    const double2 t2 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), BB(b, i)), AA(j, m)),
                                            mat(I, l2(a, n), l2(i, b)));
                    const double2 t3 = cmul(cmul(cmul(cmul(BB(n, a), BB(a, b)), AA(b, i)), AA(j, m)),
                                            mat(K, l2(i, b), l2(n, a)));
                    const double2 t4 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, i)), BB(j, b)), AA(b, m)),
                                            X[l6(a, n, j, b, n, i)]);
                    const double2 t5 = cmul(cmul(cmul(cmul(BB(n, a), AA(a, i)), BB(j, b)), AA(b, m)),
                                            X[l6(i, a, j, b, n, i)]);
                    const double2 t6 = cmul(cmul(cmul(cmul(AA(n, a), AA(a, b)), BB(b, i)), BB(j, m)),
                                            X[l6(b, a, j, m, a, i)]);
                    const double2 t7 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), AA(b, i)), BB(j, m)),
                                            X[l6(i, b, j, m, a, i)]);
                    res = csub(res, cadd(csub(csub(cadd(cadd(csub(t1, t2), t3), t4), t5), t6), t7));
                }
            if (j == m)
                for (int a = 0; a < N; ++a)
                    for (int b = 0; b < N; ++b)
                        for (int c = 0; c < N; ++c) {
                            const double2 add1 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), AA(b, c)), BB(c, i)),
                                                      mat(M, l2(c, i), l2(a, b)));
                            const double2 add2 = cmul(cmul(cmul(cmul(AA(n, a), BB(a, b)), BB(b, c)), AA(c, i)),
                                                      mat(K, l2(i, c), l2(a, b)));
                            const double2 nA = make_double2(-AA(n, a).x, -AA(n, a).y);
                            const double2 add3 = cmul(cmul(cmul(cmul(nA, AA(a, b)), BB(b, c)), BB(c, i)),
                                                      mat(I, l2(b, a), l2(i, c)));
                            res = cadd(res, cadd(cadd(add1, add2), add3));
                        }
#undef AA
#undef BB
        }
    }
    GW[idx] = res;
}

// GW = T + T^H in place, one thread per unordered pair (r <= c).
__global__ __launch_bounds__(kBlock) void k_gw_sym(double2* GW, int N2) {
    const long e = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x;
    const long npairs = static_cast<long>(N2) * (N2 + 1) / 2;
    if (e >= npairs) return;
    // pair e -> (r, c) with c the column: e = c (c + 1) / 2 + r, r <= c
    long c = static_cast<long>((sqrt(8.0 * static_cast<double>(e) + 1.0) - 1.0) / 2.0);
    while (c * (c + 1) / 2 > e) --c;
    while ((c + 1) * (c + 2) / 2 <= e) ++c;
    const long r = e - c * (c + 1) / 2;
    const double2 g = cadd(GW[r + c * N2], cconj(GW[c + r * N2]));
    GW[r + c * N2] = g;
    if (r != c) GW[c + r * N2] = cconj(g);
}

// ---------------------------------------------------------------------------
// Factorised L4 for any N (used for N > 8, where k_l4's N^4 LDS tile does
// not fit).  Every NAKZWAN term of T(n,i;m,j) (tcl4_assemble.cpp:29-178)
// factors through quantities of (n,i) and a bucket index:
//   -t1 + t2 - t3 = A(j,m) S123(n,i)
//   t6 - t7       = B(j,m) S67(n,i, f(j,m)),  S67 = sum_a A(n,a) Yd(a,i,.)
//   t4 - t5       = sum_b B(j,b) A(b,m) Wd(n,i, f(j,b))
//   add1..3       = d(j,m) D(n,i),            D = sum_a A(n,a) E(a,i)
// with Yd(a,i,beta) = sum_b [A(a,b)B(b,i) X[f(b,a)][beta][f(a,i)]
//                            - B(a,b)A(b,i) X[f(i,b)][beta][f(a,i)]],
// Wd(n,i,g) = sum_a [A(n,a)B(a,i) X[f(a,n)][g][f(n,i)]
//                    - B(n,a)A(a,i) X[f(i,a)][g][f(n,i)]],  X = C + R,
// so T costs O(N^5 + N^3 nf) per time instead of O(N^6).  The sums run in a
// different order from the reference's, so values agree to rounding, not
// bit for bit.  One workgroup per time; the intermediates of a time live in
// its slice of a global scratch (L2-resident).
// ---------------------------------------------------------------------------
struct L4fArgs {
    const double2* Fs;   // [npos][nf3]
    const double2* Cs;
    const double2* Rs;
    const int* p2f;      // N x N column-major
    const double2* ops;  // num_ops x N x N
    int N, nf, num_ops;
    long nf3;
    double2* L4;         // [npos][N^4]
    double2* GW;         // optional [npos][N^4]
    double2* scr;        // [npos][scr_per]
    long scr_per;        // complex values per position
};

template <int BS>
__global__ __launch_bounds__(BS) void k_l4f(L4fArgs A) {
    extern __shared__ double2 sm[];
    const int N = A.N, N2 = N * N, nf = A.nf, P = A.num_ops * A.num_ops;
    const long pos = blockIdx.x;
    double2* opsS = sm;                                          // num_ops * N^2
    int* f = reinterpret_cast<int*>(opsS + A.num_ops * N2);      // N^2
    for (int e = threadIdx.x; e < A.num_ops * N2; e += BS) opsS[e] = A.ops[e];
    int bad = 0;
    for (int e = threadIdx.x; e < N2; e += BS) {
        f[e] = A.p2f[e];
        bad |= (f[e] < 0 || f[e] >= nf);
    }
    // A bucket index outside [0, nf) (a caller's device map that the host
    // never saw) would index outside the slices: the time's L4 becomes NaN.
    if (__syncthreads_or(bad)) {
        const double nan = __builtin_nan("");
        const long N4 = static_cast<long>(N2) * N2;
        for (long idx = threadIdx.x; idx < N4; idx += BS) {
            A.L4[pos * N4 + idx] = make_double2(nan, nan);
            if (A.GW) A.GW[pos * N4 + idx] = make_double2(nan, nan);
        }
        return;
    }
    const double2* Fp = A.Fs + pos * A.nf3;
    const double2* Cp = A.Cs + pos * A.nf3;
    const double2* Rp = A.Rs + pos * A.nf3;
    const long nfl = nf;
    auto id3 = [nfl](int x, int y, int z) { return (static_cast<long>(x) * nfl + y) * nfl + z; };
    auto F2 = [f, N](int r, int c) { return f[r + c * N]; };
    auto X = [&](long e) { return cadd(Cp[e], Rp[e]); };
    // scratch of this time: per pair [S123 N2][D N2][S67 N2 nf][Wd N2 nf], then Yd N2 nf and E N2
    double2* base = A.scr + pos * A.scr_per;
    const long per_pair = 2L * N2 + 2L * N2 * nf;
    double2* Yd = base + P * per_pair;
    double2* Ev = Yd + static_cast<long>(N2) * nf;
    for (int iA = 0; iA < A.num_ops; ++iA)
        for (int iB = 0; iB < A.num_ops; ++iB) {
            const double2* Ap = opsS + iA * N2;
            const double2* Bp = opsS + iB * N2;
#define AA(r, c) Ap[(r) + (c) * N]
#define BB(r, c) Bp[(r) + (c) * N]
            double2* pb = base + (iA * A.num_ops + iB) * per_pair;
            double2* S123 = pb;
            double2* Dv = pb + N2;
            double2* S67 = pb + 2 * N2;
            double2* Wd = S67 + static_cast<long>(N2) * nf;
            // stage 1: Yd(a,i,beta), Wd(n,i,g), S123(n,i), E(a,i)
            for (long e = threadIdx.x; e < static_cast<long>(N2) * nf; e += BS) {
                const int a = static_cast<int>(e % N), i = static_cast<int>((e / N) % N);
                const int g = static_cast<int>(e / N2);
                double2 y = make_double2(0.0, 0.0), w = y;
                for (int b = 0; b < N; ++b) {
                    y = cadd(y, cmul(cmul(AA(a, b), BB(b, i)), X(id3(F2(b, a), g, F2(a, i)))));
                    y = csub(y, cmul(cmul(BB(a, b), AA(b, i)), X(id3(F2(i, b), g, F2(a, i)))));
                }
                // Wd uses (n, i) = (a, i) of this index
                for (int c = 0; c < N; ++c) {
                    w = cadd(w, cmul(cmul(AA(a, c), BB(c, i)), X(id3(F2(c, a), g, F2(a, i)))));
                    w = csub(w, cmul(cmul(BB(a, c), AA(c, i)), X(id3(F2(i, c), g, F2(a, i)))));
                }
                Yd[e] = y;
                Wd[e] = w;
            }
            for (int e = threadIdx.x; e < N2; e += BS) {
                const int n = e % N, i = e / N;
                double2 s = make_double2(0.0, 0.0), ev = s;
                for (int a = 0; a < N; ++a)
                    for (int b = 0; b < N; ++b) {
                        const double2 Mv = csub(Fp[id3(F2(b, a), F2(b, i), F2(n, b))], Rp[id3(F2(b, a), F2(n, a), F2(a, i))]);
                        const double2 Iv = Fp[id3(F2(a, n), F2(b, i), F2(n, b))];
                        const double2 Kv = Rp[id3(F2(i, b), F2(n, a), F2(a, i))];
                        s = csub(s, cmul(cmul(cmul(BB(n, a), AA(a, b)), BB(b, i)), Mv));
                        s = cadd(s, cmul(cmul(cmul(AA(n, a), BB(a, b)), BB(b, i)), Iv));
                        s = csub(s, cmul(cmul(cmul(BB(n, a), BB(a, b)), AA(b, i)), Kv));
                        // E(a', i) with a' = n: sum over (b, c) = (a, b) of this loop
                        const int ap = n, bb = a, cc = b;
                        const double2 M2 = csub(Fp[id3(F2(cc, bb), F2(cc, i), F2(ap, cc))],
                                                Rp[id3(F2(cc, bb), F2(ap, bb), F2(bb, i))]);
                        const double2 K2 = Rp[id3(F2(i, cc), F2(ap, bb), F2(bb, i))];
                        const double2 I2 = Fp[id3(F2(bb, ap), F2(cc, i), F2(ap, cc))];
                        ev = cadd(ev, cmul(cmul(cmul(BB(ap, bb), AA(bb, cc)), BB(cc, i)), M2));
                        ev = cadd(ev, cmul(cmul(cmul(BB(ap, bb), BB(bb, cc)), AA(cc, i)), K2));
                        ev = csub(ev, cmul(cmul(cmul(AA(ap, bb), BB(bb, cc)), BB(cc, i)), I2));
                    }
                S123[e] = s;
                Ev[e] = ev;
            }
            __syncthreads();
            // stage 2: S67(n,i,beta) = sum_a A(n,a) Yd(a,i,beta); D(n,i) = sum_a A(n,a) E(a,i)
            for (long e = threadIdx.x; e < static_cast<long>(N2) * nf; e += BS) {
                const int n = static_cast<int>(e % N);
                const long rest = e - n;  // (i, beta) part of the index
                double2 s = make_double2(0.0, 0.0);
                for (int a = 0; a < N; ++a) s = cadd(s, cmul(AA(n, a), Yd[rest + a]));
                S67[e] = s;
            }
            for (int e = threadIdx.x; e < N2; e += BS) {
                const int n = e % N, i = e / N;
                double2 s = make_double2(0.0, 0.0);
                for (int a = 0; a < N; ++a) s = cadd(s, cmul(AA(n, a), Ev[a + i * N]));
                Dv[e] = s;
            }
            __syncthreads();
#undef AA
#undef BB
        }
    // stage 3: T(n,i;m,j) from the intermediates; L4(n,m;i,j) = T(n,i;m,j) + conj(T(m,j;n,i))
    auto Tval = [&](int n, int i, int m, int j) {
        double2 t = make_double2(0.0, 0.0);
        const int ni = n + i * N;
        for (int iA = 0; iA < A.num_ops; ++iA)
            for (int iB = 0; iB < A.num_ops; ++iB) {
                const double2* Ap = opsS + iA * N2;
                const double2* Bp = opsS + iB * N2;
                const double2* pb = base + (iA * A.num_ops + iB) * per_pair;
                const double2* S67 = pb + 2 * N2;
                const double2* Wd = S67 + static_cast<long>(N2) * nf;
                t = cadd(t, cmul(Ap[j + m * N], pb[ni]));
                t = cadd(t, cmul(Bp[j + m * N], S67[ni + static_cast<long>(F2(j, m)) * N2]));
                for (int b = 0; b < N; ++b)
                    t = csub(t, cmul(cmul(Bp[j + b * N], Ap[b + m * N]), Wd[ni + static_cast<long>(F2(j, b)) * N2]));
                if (j == m) t = cadd(t, pb[N2 + ni]);
            }
        return t;
    };
    const long N4 = static_cast<long>(N2) * N2;
    double2* L = A.L4 + pos * N4;
    for (long idx = threadIdx.x; idx < N4; idx += BS) {
        const int r = static_cast<int>(idx % N2), c = static_cast<int>(idx / N2);  // r = (n, i), c = (m, j)
        const int n = r % N, i = r / N, m = c % N, j = c / N;
        const double2 g = cadd(Tval(n, i, m, j), cconj(Tval(m, j, n, i)));
        if (A.GW) A.GW[pos * N4 + idx] = g;
        L[(n + m * N) + static_cast<long>(i + j * N) * N2] = g;
    }
}

// Per-time workgroups of the forward L4 and TCL2 kernels: every entry is
// formed by one thread, so the size only decides which thread owns it (the
// results are identical at any size).  Large N (d >= 12: one workgroup per
// time over tens of thousands of entries) runs 1024 threads, whose sixteen
// waves per CU hide the scattered slice and Gamma reads; small N keeps 256.
constexpr int kFwdBlockLarge = 1024;
constexpr int kFwdLargeN = 12;
static void launch_k_l4f(unsigned grid, int N, std::size_t lds, hipStream_t st, const L4fArgs& F) {
    if (N >= kFwdLargeN)
        hipLaunchKernelGGL(k_l4f<kFwdBlockLarge>, dim3(grid), dim3(kFwdBlockLarge), lds, st, F);
    else
        hipLaunchKernelGGL(k_l4f<kBlock>, dim3(grid), dim3(kBlock), lds, st, F);
}

// ---------------------------------------------------------------------------
// RK4 on vec(rho) with prebuilt generators (rk4_dense.cpp:133-160, 246-320),
// one workgroup, D <= 1024.  Matvecs sum over j in order per row.
// ---------------------------------------------------------------------------
// Matrices are read as L[r + c * ld]: ld = D in global memory, D + 1 when
// staged in LDS (the padded column stride spreads the column reads of L^H
// products over the banks).  Everything down to the kernels is inlined and
// each code path sees either LDS or global matrices, so the compiler knows
// the address space of every access (LDS reads, not flat ones).  Every product sums over c in order per row, so
// staged and unstaged runs, and every instantiation below, give identical
// values.  DT > 0 fixes D at compile time (the row loop is unrolled, so its
// loads issue together); DT = 0 is the general case.
// In the general case (DT = 0, D a runtime value) the columns are taken in
// blocks of kMvBlk whose loads are issued before the block's products, so a
// thread keeps kMvBlk generator loads in flight (the sum order is unchanged).
constexpr int kMvBlk = 16;

template <int DT>
__device__ __forceinline__ void matvec_rows(const double2* __restrict__ L, int ld, const double2* x, double2* y,
                                            int D) {
    const int n = DT ? DT : D;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double2 s = make_double2(0.0, 0.0);
        if constexpr (DT == 0) {
            int c = 0;
            for (; c + kMvBlk <= n; c += kMvBlk) {
                double2 v[kMvBlk];
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) v[j] = L[r + static_cast<long>(c + j) * ld];
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) s = cadd(s, cmul(v[j], x[c + j]));
            }
            for (; c < n; ++c) s = cadd(s, cmul(L[r + static_cast<long>(c) * ld], x[c]));
        } else {
#pragma unroll
            for (int c = 0; c < DT; ++c) s = cadd(s, cmul(L[r + static_cast<long>(c) * ld], x[c]));
        }
        y[r] = s;
    }
}

// y = 0.5 (L0 + L1) x, the midpoint generator formed per element.
template <int DT>
__device__ __forceinline__ void matvec_half(const double2* __restrict__ L0, const double2* __restrict__ L1, int ld,
                                            const double2* x, double2* y, int D) {
    const int n = DT ? DT : D;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double2 s = make_double2(0.0, 0.0);
        if constexpr (DT == 0) {
            int c = 0;
            for (; c + kMvBlk <= n; c += kMvBlk) {
                double2 v0[kMvBlk], v1[kMvBlk];
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) {
                    const long e = r + static_cast<long>(c + j) * ld;
                    v0[j] = L0[e];
                    v1[j] = L1[e];
                }
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) s = cadd(s, cmul(cscale(cadd(v0[j], v1[j]), 0.5), x[c + j]));
            }
            for (; c < n; ++c) {
                const long e = r + static_cast<long>(c) * ld;
                s = cadd(s, cmul(cscale(cadd(L0[e], L1[e]), 0.5), x[c]));
            }
        } else {
#pragma unroll
            for (int c = 0; c < DT; ++c) {
                const long e = r + static_cast<long>(c) * ld;
                const double2 h = cscale(cadd(L0[e], L1[e]), 0.5);
                s = cadd(s, cmul(h, x[c]));
            }
        }
        y[r] = s;
    }
}

// One RK4 step x <- x + dt/6 (k1 + 2 k2 + 2 k3 + k4) by the whole workgroup
// (rk4_dense.cpp:133-160); Lh null: midpoint 0.5 (L0 + L1) per element.
// L0 and L1 have column stride ld; Lh (global) has D.
template <int DT>
__device__ __forceinline__ void rk4_step(const double2* L0, const double2* L1, int ld, const double2* Lh, double2* x,
                                         double2* k1, double2* k2, double2* k3, double2* k4, double2* tmp, int D,
                                         double dt, double2* stash = nullptr) {
    // stash (optional): receives the stage points u1, u2, u3 (3 x D)
    const double hdt = 0.5 * dt;
    const double s6 = dt / 6.0;
    matvec_rows<DT>(L0, ld, x, k1, D);
    __syncthreads();
    for (int e = threadIdx.x; e < D; e += blockDim.x) {
        tmp[e] = cadd(x[e], cscale(k1[e], hdt));
        if (stash) stash[e] = tmp[e];
    }
    __syncthreads();
    if (Lh) matvec_rows<DT>(Lh, D, tmp, k2, D); else matvec_half<DT>(L0, L1, ld, tmp, k2, D);
    __syncthreads();
    for (int e = threadIdx.x; e < D; e += blockDim.x) {
        tmp[e] = cadd(x[e], cscale(k2[e], hdt));
        if (stash) stash[D + e] = tmp[e];
    }
    __syncthreads();
    if (Lh) matvec_rows<DT>(Lh, D, tmp, k3, D); else matvec_half<DT>(L0, L1, ld, tmp, k3, D);
    __syncthreads();
    for (int e = threadIdx.x; e < D; e += blockDim.x) {
        tmp[e] = cadd(x[e], cscale(k3[e], dt));
        if (stash) stash[2 * D + e] = tmp[e];
    }
    __syncthreads();
    matvec_rows<DT>(L1, ld, tmp, k4, D);
    __syncthreads();
    for (int e = threadIdx.x; e < D; e += blockDim.x) {
        const double2 acc = cadd(cadd(cadd(k1[e], cscale(k2[e], 2.0)), cscale(k3[e], 2.0)), k4[e]);
        x[e] = cadd(x[e], cscale(acc, s6));
    }
    __syncthreads();
}

// Copies generators L_first .. L_first+count-1 into LDS with column stride D + 1.
__device__ __forceinline__ void stage_mats(double2* dst, const double2* Ls, long first, int count, int D) {
    const long D2 = static_cast<long>(D) * D;
    const int ld = D + 1;
    const double2* src = Ls + first * D2;
    for (long e = threadIdx.x; e < count * D2; e += blockDim.x) {
        const long m = e / D2, w = e % D2;
        const int r = static_cast<int>(w % D), c = static_cast<int>(w / D);
        dst[m * D * ld + r + static_cast<long>(c) * ld] = src[e];
    }
}

// One system's RK4 run by the whole workgroup; x, k1..k4, tmp in LDS, then
// (nstage >= 2, no Lh) the generators staged nstage at a time.  Reads the
// initial vector from r_in and writes the final one to r_out.
template <int DT>
__device__ __forceinline__ void rk4_body(double2* sh, int nstage, const double2* Ls, const double2* Lh, long steps, int D,
                         const double2* r_in, double2* r_out, double dt, long every, double2* samples,
                         long* nsamples) {
    double2* x = sh;
    double2* k1 = x + D;
    double2* k2 = k1 + D;
    double2* k3 = k2 + D;
    double2* k4 = k3 + D;
    double2* tmp = k4 + D;
    double2* M = tmp + D;
    for (int e = threadIdx.x; e < D; e += blockDim.x) x[e] = r_in[e];
    __syncthreads();
    const long D2 = static_cast<long>(D) * D;
    const int ld = D + 1;
    long ns = 0;
    if (nstage >= 2 && !Lh) {
        for (long s = 0; s < steps; s += nstage - 1) {
            const int cnt = static_cast<int>(min(static_cast<long>(nstage), steps - s + 1));
            stage_mats(M, Ls, s, cnt, D);
            __syncthreads();
            for (int q = 0; q + 1 < cnt; ++q) {
                const long st = s + q;
                if (samples && st % every == 0) {
                    for (int e = threadIdx.x; e < D; e += blockDim.x) samples[ns * D + e] = x[e];
                    ++ns;
                }
                rk4_step<DT>(M + static_cast<long>(q) * D * ld, M + static_cast<long>(q + 1) * D * ld, ld, nullptr,
                             x, k1, k2, k3, k4, tmp, D, dt);
            }
        }
    } else {
        for (long st = 0; st < steps; ++st) {
            if (samples && st % every == 0) {
                for (int e = threadIdx.x; e < D; e += blockDim.x) samples[ns * D + e] = x[e];
                ++ns;
            }
            rk4_step<DT>(Ls + st * D2, Ls + (st + 1) * D2, D, Lh ? Lh + st * D2 : nullptr, x, k1, k2, k3, k4, tmp,
                         D, dt);
        }
    }
    if (samples) {
        for (int e = threadIdx.x; e < D; e += blockDim.x) samples[ns * D + e] = x[e];
        ++ns;
    }
    for (int e = threadIdx.x; e < D; e += blockDim.x) r_out[e] = x[e];
    if (threadIdx.x == 0 && nsamples) *nsamples = ns;
}

// Dispatch on D = N^2 for the sizes of N = 2..8 (unrolled rows), else general.
#define TACO_RK4_DISPATCH(D, CALL)      \
    switch (D) {                        \
        case 4: CALL(4); break;         \
        case 9: CALL(9); break;         \
        case 16: CALL(16); break;       \
        case 25: CALL(25); break;       \
        case 36: CALL(36); break;       \
        case 49: CALL(49); break;       \
        case 64: CALL(64); break;       \
        default: CALL(0); break;        \
    }

#include "rk4_rows.hpp"

// Largest RK4 workgroup: the row-group path at D = 64.
constexpr int kRkBlockMax = 512;

// Row-group path for the sizes that have one (no separate midpoint series),
// else the general path above.
template <int DT>
__device__ __forceinline__ void rk4_any(double2* sh, int nstage, const double2* Ls, const double2* Lh, long steps,
                                        int D, const double2* r_in, double2* r_out, double dt, long every,
                                        double2* samples, long* nsamples) {
    if constexpr (DT != 0) {
        if (!Lh) {
            rk4_rows<DT, false>(sh, Ls, 0, steps, r_in, r_out, dt, every, samples, nsamples, nullptr);
            return;
        }
    }
    rk4_body<DT>(sh, nstage, Ls, Lh, steps, D, r_in, r_out, dt, every, samples, nsamples);
}

// One system on the row-group path (D = N^2, N = 2..8), 64 W threads.
template <int DT>
__global__ __launch_bounds__(64 * RowGroup<DT>::W) void k_rk4_rows(const double2* Ls, long steps, double2* r,
                                                                   double dt, long every, double2* samples,
                                                                   long* nsamples) {
    extern __shared__ double2 sh[];
    rk4_rows<DT, false>(sh, Ls, 0, steps, r, r, dt, every, samples, nsamples, nullptr);
}

// Calls f(std::integral_constant<int, D>) for a D with a row-group path.
template <class F>
static bool rows_dispatch(int D, F&& f) {
    switch (D) {
        case 4: f(std::integral_constant<int, 4>{}); return true;
        case 9: f(std::integral_constant<int, 9>{}); return true;
        case 16: f(std::integral_constant<int, 16>{}); return true;
        case 25: f(std::integral_constant<int, 25>{}); return true;
        case 36: f(std::integral_constant<int, 36>{}); return true;
        case 49: f(std::integral_constant<int, 49>{}); return true;
        case 64: f(std::integral_constant<int, 64>{}); return true;
        default: return false;
    }
}

__global__ __launch_bounds__(kBlock) void k_rk4(const double2* Ls, const double2* Lh, long steps, int D, double2* r,
                                                double dt, long every, double2* samples, long* nsamples, int nstage) {
    extern __shared__ double2 sh[];
#define TACO_RK4_CALL(DT) rk4_body<DT>(sh, nstage, Ls, Lh, steps, D, r, r, dt, every, samples, nsamples)
    TACO_RK4_DISPATCH(D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
}

// Many independent systems, one workgroup each, in a single launch.
struct Rk4Sys {
    const double2* Ls;   // [steps + 1][D x D]
    int nstage;          // generators staged in LDS at a time (0: read from global)
    const double2* r0;   // [D]
    double2* r;          // [D] final vector
    double2* samples;    // [steps + 1][D]
    long steps;
    double dt;
    int D;
    const double2* Lh = nullptr;  // [steps][D x D] midpoints (large D), else formed on the fly
};

__global__ __launch_bounds__(kRkBlockMax) void k_rk4_batch(const Rk4Sys* sys) {
    extern __shared__ double2 sh[];
    const Rk4Sys S = sys[blockIdx.x];
#define TACO_RK4_CALL(DT) rk4_any<DT>(sh, S.nstage, S.Ls, S.Lh, S.steps, S.D, S.r0, S.r, S.dt, 1, S.samples, nullptr)
    TACO_RK4_DISPATCH(S.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
}

// dst[q] = src[rows[q]] for rows of `len` complex values.
__global__ __launch_bounds__(kBlock) void k_gather_rows(const double2* src, const int* rows, long len, double2* dst) {
    const double2* s = src + static_cast<long>(rows[blockIdx.x]) * len;
    double2* d = dst + static_cast<long>(blockIdx.x) * len;
    for (long e = threadIdx.x; e < len; e += kBlock) d[e] = s[e];
}

// out[n] = (L[n] + L[n + 1]) / 2 elementwise for n < count (midpoint
// generators, formed with the expression the RK4 products use on the fly).
__global__ __launch_bounds__(kBlock) void k_midpoints(const double2* __restrict__ L, long count, long D2,
                                                      double2* __restrict__ out) {
    const long total = count * D2;
    for (long i = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long>(gridDim.x) * kBlock)
        out[i] = cscale(cadd(L[i], L[i + D2]), 0.5);
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(double2* dst, long count) {
    const long i = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x;
    if (i < count) dst[i] = make_double2(0.0, 0.0);
}

// TCL2 generator at every t (generator.cpp:81-177 with the drivers' single-
// or diagonal-channel kernels) added onto L_all[t] (which holds L4(t) or 0).
// One workgroup per time; each thread builds whole columns.
struct Tcl2Args {
    const double2* gamma;  // [nf][Nt]
    long Nt;
    const double* eps;
    const double2* ops;    // A_eig
    const int* p2f;
    int N, nf, num_ops;
    double2* L;            // [T][D*D], accumulated into
    int order;             // 0: unitary only; 2/4: full TCL2
    int use_lists = 1;     // bucket-list path for N <= kTcl2ListN (TACO_TCL2_LISTS=0: all-bucket walk)
    // Channel matrices (generator.cpp:92-162): when non-null, Gamma_b(alpha, beta)
    // at column (b num_ops + beta) num_ops + alpha, time contiguous ([cols][Nt]),
    // replaces the diagonal gamma; skip[col] marks the (b, alpha, beta) whose
    // A_alpha,b^dag A_beta,b is the zero matrix (the reference skips them).
    const double2* gamma_ab = nullptr;
    const unsigned char* skip = nullptr;
    const long* tmap = nullptr;  // block -> Gamma row (null: the block index)
};

// Dynamic LDS of k_tcl2: the general path keeps H (N^2); the bucket-list path
// (N <= kTcl2ListN) adds the pair map, and per ordered pair (x, y) the
// ascending list of buckets b with some q such that p2f(q,x) == p2f(q,y) == b.
constexpr int kTcl2ListN = 16;
__host__ __device__ inline std::size_t tcl2_lds(int N) {
    const std::size_t N2 = static_cast<std::size_t>(N) * N;
    if (N > kTcl2ListN) return 2 * N2 * 16;
    return N2 * 16 + N2 * 4 + N2 * 4 + N2 * static_cast<std::size_t>(N) * 2;
}

// L_TCL2(t) from channel matrices Gamma_b(alpha, beta) (build_tcl2_components,
// generator.cpp:81-177, in its general form): every entry walks (b, alpha,
// beta) in the reference's order; a (b, alpha, beta) with A_alpha,b^dag
// A_beta,b == 0, or with a zero jump rate 2 Re Gamma / Lamb-shift rate Im Gamma,
// contributes nothing to the respective part, as there.
template <int BS>
__device__ void tcl2_channels(const Tcl2Args& A, long t, double2* H, double2* L) {
    const int N = A.N, N2 = N * N, D = N2, no = A.num_ops;
    auto G = [&](int b, int al, int be) {
        return A.gamma_ab[(static_cast<long>(b * no + be) * no + al) * A.Nt + t];
    };
    auto skipped = [&](int b, int al, int be) { return A.skip[(b * no + be) * no + al] != 0; };
    // (A_al,b^dag A_be,b)(r, c) = sum_q [p2f(q,r) == p2f(q,c) == b] conj(A_al(q,r)) A_be(q,c)
    auto adag_a = [&](const double2* Aa, const double2* Ab, int r, int c, int b) {
        double2 s = make_double2(0, 0);
        for (int q = 0; q < N; ++q)
            if (A.p2f[q + r * N] == b && A.p2f[q + c * N] == b) s = cadd(s, cmul(cconj(Aa[q + r * N]), Ab[q + c * N]));
        return s;
    };
    for (int e = threadIdx.x; e < N2; e += BS) {
        const int r = e % N, c = e / N;
        double2 h = make_double2(0.0, 0.0), hc = make_double2(0.0, 0.0);
        if (A.order != 0)
            for (int b = 0; b < A.nf; ++b)
                for (int al = 0; al < no; ++al)
                    for (int be = 0; be < no; ++be) {
                        if (skipped(b, al, be)) continue;
                        const double S = G(b, al, be).y;
                        if (S == 0.0) continue;
                        const double2* Aa = A.ops + al * N2;
                        const double2* Ab = A.ops + be * N2;
                        h = cadd(h, cscale(adag_a(Aa, Ab, r, c, b), S));
                        hc = cadd(hc, cscale(adag_a(Aa, Ab, c, r, b), S));
                    }
        double2 v = cscale(cadd(h, cconj(hc)), 0.5);
        if (r == c) v.x += A.eps[r];
        H[e] = v;
    }
    __syncthreads();
    for (long e = threadIdx.x; e < static_cast<long>(D) * D; e += BS) {
        const int row = static_cast<int>(e % D), col = static_cast<int>(e / D);
        const int r1 = row / N, r2 = row % N, c1 = col / N, c2 = col % N;
        const double2 left = (r1 == c1) ? H[r2 + c2 * N] : make_double2(0, 0);
        const double2 right = (r2 == c2) ? H[c1 + r1 * N] : make_double2(0, 0);
        const double2 d = csub(left, right);
        double2 diss = make_double2(0.0, 0.0);
        if (A.order != 0)
            for (int b = 0; b < A.nf; ++b)
                for (int al = 0; al < no; ++al)
                    for (int be = 0; be < no; ++be) {
                        if (skipped(b, al, be)) continue;
                        const double gam = 2.0 * G(b, al, be).x;
                        if (gam == 0.0) continue;
                        const double2* Aa = A.ops + al * N2;
                        const double2* Ab = A.ops + be * N2;
                        double2 jump = make_double2(0, 0);
                        if (A.p2f[r1 + c1 * N] == b && A.p2f[r2 + c2 * N] == b)
                            jump = cmul(cconj(Aa[r1 + c1 * N]), Ab[r2 + c2 * N]);
                        const double2 lt = (r1 == c1) ? adag_a(Aa, Ab, r2, c2, b) : make_double2(0, 0);
                        const double2 rt = (r2 == c2) ? adag_a(Aa, Ab, c1, r1, b) : make_double2(0, 0);
                        diss = cadd(diss, cscale(csub(csub(jump, cscale(lt, 0.5)), cscale(rt, 0.5)), gam));
                    }
        L[e] = cadd(L[e], cadd(make_double2(d.y, -d.x), diss));  // -i d + L_diss
    }
}

// L_TCL2(t) added onto L(t), one workgroup per time (generator.cpp:81-177 as
// composed in TCL4_spin_boson_example.cpp:303-322).  Every entry sums its
// bucket terms in ascending bucket order; a bucket whose term is exactly zero
// (no jump, Lamb-shift or anticommutator element of the entry in it) adds
// nothing, so for N <= 16 each entry visits only the buckets of its pair
// lists: one bucket at most off the N-diagonal blocks, the merged lists on
// them.  The values equal the all-bucket walk (general path) up to the sign
// of an exact zero.
template <int BS>
__global__ __launch_bounds__(BS) void k_tcl2(Tcl2Args A) {
    extern __shared__ double2 sh[];
    const int N = A.N, N2 = N * N, D = N2;
    const long t = A.tmap ? A.tmap[blockIdx.x] : static_cast<long>(blockIdx.x);
    double2* H = sh;  // N x N effective Hamiltonian
    double2* L = A.L + static_cast<long>(blockIdx.x) * D * D;
    if (A.gamma_ab) {
        tcl2_channels<BS>(A, t, H, L);
        return;
    }
    const bool lists = A.use_lists && N <= kTcl2ListN;
    int* P = reinterpret_cast<int*>(H + N2);                        // pair map
    int* cnt = P + N2;                                              // list lengths per pair
    unsigned short* bk = reinterpret_cast<unsigned short*>(cnt + N2);  // [pair][N] buckets
    if (lists) {
        for (int e = threadIdx.x; e < N2; e += BS) P[e] = A.p2f[e];
        __syncthreads();
        for (int pr = threadIdx.x; pr < N2; pr += BS) {
            const int x = pr % N, y = pr / N;
            unsigned short* l = bk + pr * N;
            int n = 0;
            for (int q = 0; q < N; ++q) {
                const int b = P[q + x * N];
                if (b < 0 || b >= A.nf || b != P[q + y * N]) continue;
                int i = n;  // insertion into the ascending list, unique
                while (i > 0 && l[i - 1] > b) --i;
                if (i > 0 && l[i - 1] == b) continue;
                for (int m = n; m > i; --m) l[m] = l[m - 1];
                l[i] = static_cast<unsigned short>(b);
                ++n;
            }
            cnt[pr] = n;
        }
        __syncthreads();
    }
    // sum_q [p2f(q,x) == p2f(q,y) == b] conj(A(q,x)) A(q,y), q ascending
    auto qsum = [&](const double2* Ap, int x, int y, int b) {
        double2 s = make_double2(0, 0);
        for (int q = 0; q < N; ++q)
            if (P[q + x * N] == b && P[q + y * N] == b) s = cadd(s, cmul(cconj(Ap[q + x * N]), Ap[q + y * N]));
        return s;
    };
    // H_eff = diag(eps) + hermitize(sum_b sum_a S_b A_b^dag A_b) with S_b = Im gamma(t,b).
    for (int e = threadIdx.x; e < N2; e += BS) {
        const int r = e % N, c = e / N;
        double2 h = make_double2(0.0, 0.0), hc = make_double2(0.0, 0.0);
        if (A.order != 0) {
            for (int al = 0; al < A.num_ops; ++al) {
                const double2* Ap = A.ops + al * N2;
                if (lists) {
                    // (A_b^dag A_b)(r,c) is nonzero only for the buckets of pair (r, c)
                    const unsigned short* l = bk + (r + c * N) * N;
                    const int n = cnt[r + c * N];
                    for (int i = 0; i < n; ++i) {
                        const int b = l[i];
                        const double S = A.gamma[static_cast<long>(b) * A.Nt + t].y;
                        if (S == 0.0) continue;
                        h = cadd(h, cscale(qsum(Ap, r, c, b), S));
                        hc = cadd(hc, cscale(qsum(Ap, c, r, b), S));
                    }
                    continue;
                }
                for (int b = 0; b < A.nf; ++b) {
                    const double S = A.gamma[static_cast<long>(b) * A.Nt + t].y;
                    if (S == 0.0) continue;
                    // (A_b^dag A_b)(r,c) = sum_q conj(A_b(q,r)) A_b(q,c), A_b = A masked to bucket b
                    double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
                    for (int q = 0; q < N; ++q) {
                        if (A.p2f[q + r * N] == b && A.p2f[q + c * N] == b)
                            s1 = cadd(s1, cmul(cconj(Ap[q + r * N]), Ap[q + c * N]));
                        if (A.p2f[q + c * N] == b && A.p2f[q + r * N] == b)
                            s2 = cadd(s2, cmul(cconj(Ap[q + c * N]), Ap[q + r * N]));
                    }
                    h = cadd(h, cscale(s1, S));
                    hc = cadd(hc, cscale(s2, S));
                }
            }
        }
        double2 v = cscale(cadd(h, cconj(hc)), 0.5);
        if (r == c) v.x += A.eps[r];
        H[e] = v;
    }
    __syncthreads();
    // L(row=(r1,r2), col=(c1,c2)) with row = r1*N + r2 (kron(I,H) - kron(H^T,I) convention)
    const double2 zero = make_double2(0, 0);
    for (long e = threadIdx.x; e < static_cast<long>(D) * D; e += BS) {
        const int row = static_cast<int>(e % D), col = static_cast<int>(e / D);
        const int r1 = row / N, r2 = row % N, c1 = col / N, c2 = col % N;
        double2 left = (r1 == c1) ? H[r2 + c2 * N] : make_double2(0, 0);
        double2 right = (r2 == c2) ? H[c1 + r1 * N] : make_double2(0, 0);
        const double2 d = csub(left, right);
        double2 val = make_double2(d.y, -d.x);  // -i * d
        if (A.order != 0) {
            for (int al = 0; al < A.num_ops; ++al) {
                const double2* Ap = A.ops + al * N2;
                if (lists) {
                    const int b1 = P[r1 + c1 * N], b2 = P[r2 + c2 * N];
                    const int kInf = 1 << 30;
                    const int jb = (b1 == b2 && b1 >= 0 && b1 < A.nf) ? b1 : kInf;
                    // lt: pair (r2, c2) when r1 == c1; rt: pair (c1, r1) when r2 == c2
                    const unsigned short* la = bk + (r2 + c2 * N) * N;
                    const unsigned short* lb = bk + (c1 + r1 * N) * N;
                    const int na = r1 == c1 ? cnt[r2 + c2 * N] : 0, nb = r2 == c2 ? cnt[c1 + r1 * N] : 0;
                    int ia = 0, ib = 0;
                    bool jpend = jb != kInf;
                    while (true) {
                        int b = jpend ? jb : kInf;
                        if (ia < na && la[ia] < b) b = la[ia];
                        if (ib < nb && lb[ib] < b) b = lb[ib];
                        if (b == kInf) break;
                        double2 jump = zero, lt = zero, rt = zero;
                        const bool use_a = ia < na && la[ia] == b, use_b = ib < nb && lb[ib] == b;
                        const bool use_j = jpend && jb == b;
                        if (use_a) ++ia;
                        if (use_b) ++ib;
                        if (use_j) jpend = false;
                        const double gam = 2.0 * A.gamma[static_cast<long>(b) * A.Nt + t].x;
                        if (gam == 0.0) continue;
                        if (use_j) jump = cmul(cconj(Ap[r1 + c1 * N]), Ap[r2 + c2 * N]);
                        if (use_a) lt = qsum(Ap, r2, c2, b);
                        if (use_b) rt = qsum(Ap, c1, r1, b);
                        const double2 term = csub(csub(jump, cscale(lt, 0.5)), cscale(rt, 0.5));
                        val = cadd(val, cscale(term, gam));
                    }
                    continue;
                }
                for (int b = 0; b < A.nf; ++b) {
                    const double gam = 2.0 * A.gamma[static_cast<long>(b) * A.Nt + t].x;
                    if (gam == 0.0) continue;
                    const bool in1 = A.p2f[r1 + c1 * N] == b, in2 = A.p2f[r2 + c2 * N] == b;
                    double2 jump = make_double2(0, 0);
                    if (in1 && in2) jump = cmul(cconj(Ap[r1 + c1 * N]), Ap[r2 + c2 * N]);
                    double2 lt = make_double2(0, 0), rt = make_double2(0, 0);
                    if (r1 == c1)
                        for (int q = 0; q < N; ++q)
                            if (A.p2f[q + r2 * N] == b && A.p2f[q + c2 * N] == b)
                                lt = cadd(lt, cmul(cconj(Ap[q + r2 * N]), Ap[q + c2 * N]));
                    if (r2 == c2)
                        for (int q = 0; q < N; ++q)
                            if (A.p2f[q + c1 * N] == b && A.p2f[q + r1 * N] == b)
                                rt = cadd(rt, cmul(cconj(Ap[q + c1 * N]), Ap[q + r1 * N]));
                    const double2 term = csub(csub(jump, cscale(lt, 0.5)), cscale(rt, 0.5));
                    val = cadd(val, cscale(term, gam));
                }
            }
        }
        L[e] = cadd(L[e], val);
    }
}

static void launch_tcl2(unsigned grid, int N, std::size_t lds, hipStream_t st, const Tcl2Args& A) {
    if (N >= kFwdLargeN)
        hipLaunchKernelGGL(k_tcl2<kFwdBlockLarge>, dim3(grid), dim3(kFwdBlockLarge), lds, st, A);
    else
        hipLaunchKernelGGL(k_tcl2<kBlock>, dim3(grid), dim3(kBlock), lds, st, A);
}

// ---------------------------------------------------------------------------
// Context and device workspace
// ---------------------------------------------------------------------------
// Counts device reallocations of every DevBuf in the process.  A captured
// graph holds raw device addresses; a batch replays its graph only while
// this count is the one it was captured at, and captures again otherwise.
static std::atomic<unsigned long> g_dev_realloc{0};

struct DevBuf {
    void* p = nullptr;
    std::size_t bytes = 0;
    void ensure(std::size_t need) {
        if (need <= bytes) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        g_dev_realloc.fetch_add(1);
        if (need == 0) return;
        HIP_OK(hipMalloc(&p, need));
        bytes = need;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct Twiddles {
    DevBuf std_, b_, c_;
    int L = -1;
};

// Host-evaluated phase table of the TACO_PHASE_LIBM mode (rows of
// exp(+i Omega (t dt)), t < len) and the key it was built for: the bit
// patterns of its Omegas, len and dt.
struct PhaseTab {
    DevBuf tab;
    std::vector<std::uint64_t> rows;
    std::size_t len = 0;
    std::uint64_t dt_bits = 0;
    bool valid = false;
    void swap(PhaseTab& o) {
        tab.swap(o.tab);
        rows.swap(o.rows);
        std::swap(len, o.len);
        std::swap(dt_bits, o.dt_bits);
        std::swap(valid, o.valid);
    }
};

// Device buffers of one tape.  A destroyed tape hands them back to its
// context, and the next tape on the context takes them over, so repeated
// gradients run without device allocations once the sizes have settled.
struct TapeBufs {
    DevBuf gamma, omegas, mirror, p2f, ops, eps, t2pos, jobtri, jobmask, clsstart, jobrow, slices, lgen, ck, adjl;
    DevBuf bath;  // bath-sourced tapes: C(t), phase steps and the bath reverse's scratch
    PhaseTab phtab;
    std::size_t bytes() const {
        std::size_t n = phtab.tab.bytes;
        for (const DevBuf* b : {&gamma, &omegas, &mirror, &p2f, &ops, &eps, &t2pos, &jobtri, &jobmask, &clsstart,
                                &jobrow, &slices, &lgen, &ck, &adjl, &bath})
            n += b->bytes;
        return n;
    }
    void swap(TapeBufs& o) {
        for (auto [a, b] : {std::pair{&gamma, &o.gamma}, {&omegas, &o.omegas}, {&mirror, &o.mirror}, {&p2f, &o.p2f},
                            {&ops, &o.ops}, {&eps, &o.eps}, {&t2pos, &o.t2pos}, {&jobtri, &o.jobtri},
                            {&jobmask, &o.jobmask}, {&clsstart, &o.clsstart}, {&jobrow, &o.jobrow},
                            {&slices, &o.slices}, {&lgen, &o.lgen}, {&ck, &o.ck}, {&adjl, &o.adjl},
                            {&bath, &o.bath}})
            a->swap(*b);
        phtab.swap(o.phtab);
    }
};

// Buffers of a multi-set tape.  The context keeps one set that a live
// multi-set tape at a time uses and later ones reuse (tens of GB at d = 16 with
// 64 sets, whose allocation would otherwise recur on every gradient); a tape
// created while another holds it gets buffers of its own.
struct MtapeArena {
    DevBuf slices;  // the F/C/R slices of the sets within the slice budget, one after another
    DevBuf shared;  // one set's slices at a time, for the sets past the budget and packed sets
    DevBuf pkidx;   // packed sets' triple lists (F, C, R back to back per distinct bucket map)
    DevBuf samp, rv, args, mid;
};

// Reverse-mode kernels (device code, same namespace as the forward kernels).
#include "adjoint_kernels.hpp"

}  // namespace taco_engine

namespace taco_engine {
// Kernel classes timed by the optional per-launch HIP-event profiler.
enum KernelId { K_PHASE_A = 0, K_PHASE_B, K_PHASE_C, K_COMBINE, K_L4, K_SPECTRA, K_TCL2, K_RK4, K_COUNT };

struct Profiler {
    bool on = false;
    struct Pending {
        int id;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double ms[K_COUNT] = {};
    double bytes[K_COUNT] = {};
    long launches[K_COUNT] = {};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        return e;
    }
    void resolve() {
        for (auto& p : pending) {
            HIP_OK(hipEventSynchronize(p.b));
            float t = 0.f;
            HIP_OK(hipEventElapsedTime(&t, p.a, p.b));
            ms[p.id] += t;
            bytes[p.id] += p.bytes;
            launches[p.id] += 1;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() {
        resolve();
        for (int k = 0; k < K_COUNT; ++k) ms[k] = bytes[k] = 0, launches[k] = 0;
    }
    ~Profiler() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};
}  // namespace taco_engine

namespace taco_engine {
struct L4AdjLists;
}

struct taco_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    taco_engine::Profiler prof;
    std::map<int, std::array<int, 3>> fft_choice;  // per FFT log-length: tile config of phases A, B, C
    std::map<int, std::unique_ptr<taco_engine::Twiddles>> tw;
    taco_engine::DevBuf gamma, omegas, mirror, p2f, ops, spec, inter1, inter2, conv, slices, t2pos, l4, gw, misc,
        rk, eps, jobtri, clsstart, jobmask, jobrow, comb, t2scr, rkmid, bred;
    taco_engine::PhaseTab phtab;  // phase table of the last plain-path system (TACO_PHASE_LIBM)
    std::size_t l4_lds_set = 0;  // dynamic LDS limit set on the fused L4 kernel
    std::size_t slice_budget = 0;  // bytes of F/C/R slices per pass of the batch builders (0: default)
    // term lists of the factorised L4 reverse per (pair_to_freq, N, nf)
    std::vector<std::pair<std::vector<int>, std::shared_ptr<taco_engine::L4AdjLists>>> adj_lists;
    std::vector<std::unique_ptr<taco_engine::TapeBufs>> tape_pool;  // buffers of destroyed tapes
    taco_engine::MtapeArena marena;     // multi-set tape buffers, reused across tapes
    const void* marena_owner = nullptr;  // the live multi-set tape using marena
};

namespace taco_engine {

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        HIP_OK(hipGetDevice(&prev));
        if (prev != dev) HIP_OK(hipSetDevice(dev));
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

// Brackets a launch with profiler events on the context stream when enabled.
// `bytes` is the launch's algorithmic HBM traffic (compulsory reads + writes).
template <class Fn>
static void timed(taco_ctx* c, int id, double bytes, Fn&& launch) {
    if (!c->prof.on) {
        launch();
        return;
    }
    hipEvent_t a = c->prof.get(), b = c->prof.get();
    HIP_OK(hipEventRecord(a, c->stream));
    launch();
    HIP_OK(hipEventRecord(b, c->stream));
    c->prof.pending.push_back({id, a, b, bytes});
}

// Twiddle tables of length 2^L: the reference's recurrence values, or (exact)
// each entry rounded once from quad precision (the window mode).
static const Twiddles& twiddles_for(taco_ctx* c, int L, bool exact = false) {
    const int key = L + (exact ? 64 : 0);
    auto it = c->tw.find(key);
    if (it != c->tw.end()) return *it->second;
    auto t = std::make_unique<Twiddles>();
    t->L = L;
    std::vector<double> tab;
    if (exact)
        make_twiddles_exact(L, tab);
    else
        make_twiddles(L, tab);  // stage-major standard table, 2^L - 1 entries
    const int a = (L + 1) / 2, b = L - a;
    const std::size_t nstd = (std::size_t{1} << std::max(a, b)) - 1;
    // Phase B forward stages s = a+1..L for group k1 < 2^a: T_s[k1 + v 2^a], v < 2^(s-1-a).
    const std::size_t bstride = (std::size_t{1} << b) - 1;
    std::vector<double> tb(2 * (std::size_t{1} << a) * bstride + 2);
    for (std::size_t k1 = 0; k1 < (std::size_t{1} << a); ++k1)
        for (int s = a + 1; s <= L; ++s) {
            const std::size_t cnt = std::size_t{1} << (s - 1 - a);
            for (std::size_t v = 0; v < cnt; ++v) {
                const std::size_t src = ((std::size_t{1} << (s - 1)) - 1) + k1 + (v << a);
                const std::size_t dst = k1 * bstride + (cnt - 1) + v;
                tb[2 * dst] = tab[2 * src];
                tb[2 * dst + 1] = tab[2 * src + 1];
            }
        }
    // Phase C inverse stages s = b+1..L for group q < 2^b: T_s[q + v 2^b], v < 2^(s-1-b).
    const std::size_t cstride = (std::size_t{1} << a) - 1;
    std::vector<double> tc(2 * (std::size_t{1} << b) * cstride + 2);
    for (std::size_t q = 0; q < (std::size_t{1} << b); ++q)
        for (int s = b + 1; s <= L; ++s) {
            const std::size_t cnt = std::size_t{1} << (s - 1 - b);
            for (std::size_t v = 0; v < cnt; ++v) {
                const std::size_t src = ((std::size_t{1} << (s - 1)) - 1) + q + (v << b);
                const std::size_t dst = q * cstride + (cnt - 1) + v;
                tc[2 * dst] = tab[2 * src];
                tc[2 * dst + 1] = tab[2 * src + 1];
            }
        }
    t->std_.ensure(std::max<std::size_t>(nstd, 1) * 16);
    if (nstd) HIP_OK(hipMemcpy(t->std_.p, tab.data(), nstd * 16, hipMemcpyHostToDevice));
    t->b_.ensure(tb.size() * 8);
    HIP_OK(hipMemcpy(t->b_.p, tb.data(), tb.size() * 8, hipMemcpyHostToDevice));
    t->c_.ensure(tc.size() * 8);
    HIP_OK(hipMemcpy(t->c_.p, tc.data(), tc.size() * 8, hipMemcpyHostToDevice));
    auto& ref = *t;
    c->tw.emplace(key, std::move(t));
    return ref;
}

// Launch phase kernels for a tile configuration index.
enum Phase { PH_A_TRIPLE, PH_A_SPEC, PH_B_TRIPLE, PH_B_SPEC, PH_C };

template <class Cfg>
static void launch_phase_cfg(Phase ph, const FcrArgs& P, long njobs, hipStream_t st) {
    const int ngA = 1 << P.b, ngB = 1 << P.a;
    switch (ph) {
        case PH_A_TRIPLE: {  // njobs = number of Omega classes touching the chunk
            const long blocks = njobs * (ngA / tile_groups<Cfg>(P.a, ngA));
            hipLaunchKernelGGL(k_phase_a<Cfg>, dim3(static_cast<unsigned>(blocks)), dim3(Cfg::kFT), 0, st, P);
            break;
        }
        case PH_A_SPEC: {
            const long blocks = njobs * (ngA / tile_groups<Cfg>(P.a, ngA));
            hipLaunchKernelGGL(k_phase_a_spec<Cfg>, dim3(static_cast<unsigned>(blocks)), dim3(Cfg::kFT), 0, st, P);
            break;
        }
        case PH_B_TRIPLE:
        case PH_B_SPEC: {
            const long blocks = njobs * (ngB / tile_groups<Cfg>(P.b, ngB));
            hipLaunchKernelGGL(k_phase_b<Cfg>, dim3(static_cast<unsigned>(blocks)), dim3(Cfg::kFT), 0, st, P,
                               ph == PH_B_SPEC ? 1 : 0);
            break;
        }
        case PH_C: {
            const long tiles = 2 * njobs * (ngA / tile_groups<Cfg>(P.a, ngA));
            const long blocks = (tiles + 1) / 2;
            hipLaunchKernelGGL(k_phase_c<Cfg>, dim3(static_cast<unsigned>(blocks)), dim3(Cfg::kFT), 0, st, P,
                               static_cast<int>(tiles));
            break;
        }
    }
}

static void launch_phase(int cfg, Phase ph, const FcrArgs& P, long njobs, hipStream_t st) {
    switch (cfg) {
        case 0: launch_phase_cfg<Cfg8k>(ph, P, njobs, st); break;
        case 1: launch_phase_cfg<Cfg4k>(ph, P, njobs, st); break;
        case 2: launch_phase_cfg<Cfg4kW>(ph, P, njobs, st); break;
        case 3: launch_phase_cfg<Cfg2k>(ph, P, njobs, st); break;
        case 4: launch_phase_cfg<Cfg4kX>(ph, P, njobs, st); break;
        case 5: launch_phase_cfg<Cfg2kW>(ph, P, njobs, st); break;
        default: launch_phase_cfg<Cfg8kX>(ph, P, njobs, st); break;
    }
}

static int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

// Read at each launch (a test compares the two paths in one process).
static bool tcl2_lists_enabled() {
    const char* e = std::getenv("TACO_TCL2_LISTS");
    return !(e && std::string(e) == "0");
}

static std::size_t chunk_budget_bytes() {
    static std::size_t v = [] {
        const char* s = std::getenv("TACO_FCR_CHUNK_BYTES");
        return s ? static_cast<std::size_t>(std::strtoull(s, nullptr, 10)) : (std::size_t{12} << 30);
    }();
    return v;
}

// Times every tile configuration on this chunk, phase by phase in pipeline
// order (each run rewrites the same values), and records the fastest.
static void autotune(taco_ctx* c, const FcrArgs& P, long ncls, long na, long nj, std::array<int, 3>& choice) {
    const Phase phases[3] = {PH_A_TRIPLE, PH_B_TRIPLE, PH_C};
    const long units[3] = {ncls, na, nj};
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    for (int ph = 0; ph < 3; ++ph) {
        float best = 0.f;
        int best_cfg = 0;
        for (int cfg = 0; cfg < kNumFftCfg; ++cfg) {
            launch_phase(cfg, phases[ph], P, units[ph], c->stream);  // warm
            // the fastest of three runs, so one noisy launch does not decide
            float ms = 0.f;
            for (int rep = 0; rep < 3; ++rep) {
                HIP_OK(hipEventRecord(e0, c->stream));
                launch_phase(cfg, phases[ph], P, units[ph], c->stream);
                HIP_OK(hipEventRecord(e1, c->stream));
                HIP_OK(hipEventSynchronize(e1));
                float m = 0.f;
                HIP_OK(hipEventElapsedTime(&m, e0, e1));
                if (rep == 0 || m < ms) ms = m;
            }
            if (cfg == 0 || ms < best) {
                best = ms;
                best_cfg = cfg;
            }
        }
        choice[ph] = best_cfg;
        launch_phase(best_cfg, phases[ph], P, units[ph], c->stream);  // leave this phase's output in place
    }
    HIP_OK(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// Device-side F/C/R for all nf^3 triples over t < T; outputs per CombineArgs.
struct FcrPlan {
    const double2* gamma;
    std::size_t Nt;
    int nf;
    const double* omegas;    // device
    const double* h_omegas;  // host copy (groups the triples by Omega)
    const int* mirror;
    double dt;
    std::size_t T;
    const std::vector<unsigned char>* need;  // per triple: bit0 F, bit1 C, bit2 R (null: all)
    std::size_t T_fft = 0;  // window mode: rows entering the convolutions (0: T)
    int method = TACO_FCR_CONVOLUTION;  // TACO_FCR_DIRECT: the O(T^2) sums of k_fcr_direct
};

// Triples the L4 assembly reads (tcl4_mikx.cpp:101-132 gathers through the
// NAKZWAN sums, tcl4_assemble.cpp:240-304): F at the M1 and I patterns, R at
// the M2 and K patterns and wherever X is read, C wherever X is read.  Only
// these need the FFT work; every other slice entry is never touched.
static std::vector<unsigned char> needed_triples(const int* p2f, int N, int nf) {
    const long nf_ = nf;
    std::vector<unsigned char> need(static_cast<std::size_t>(nf_ * nf_ * nf_), 0);
    auto f = [&](int r, int c) { return static_cast<long>(p2f[r + c * N]); };
    auto mark = [&](long x, long y, long z, unsigned char bit) { need[static_cast<std::size_t>((x * nf_ + y) * nf_ + z)] |= bit; };
    for (int p = 0; p < N; ++p)
        for (int q = 0; q < N; ++q)
            for (int u = 0; u < N; ++u)
                for (int v = 0; v < N; ++v) {
                    // off-diagonal terms with (b, a, i, n) = (p, q, u, v)
                    mark(f(p, q), f(p, u), f(v, p), 1);  // M: F[f(b,a)][f(b,i)][f(n,b)]
                    mark(f(p, q), f(v, q), f(q, u), 4);  // M: R[f(b,a)][f(n,a)][f(a,i)]
                    mark(f(q, v), f(p, u), f(v, p), 1);  // I: F[f(a,n)][f(b,i)][f(n,b)]
                    mark(f(u, p), f(v, q), f(q, u), 4);  // K: R[f(i,b)][f(n,a)][f(a,i)]
                    // diagonal terms with (c, b, a, i) = (p, q, u, v)
                    mark(f(p, q), f(p, v), f(u, p), 1);  // M: F[f(c,b)][f(c,i)][f(a,c)]
                    mark(f(p, q), f(u, q), f(q, v), 4);  // M: R[f(c,b)][f(a,b)][f(b,i)]
                    mark(f(v, p), f(u, q), f(q, v), 4);  // K: R[f(i,c)][f(a,b)][f(b,i)]
                    mark(f(q, u), f(p, v), f(u, p), 1);  // I: F[f(b,a)][f(c,i)][f(a,c)]
                }
    // X = C + R at (f(a,n), f(j,b), f(n,i)), (f(i,a), f(j,b), f(n,i)),
    // (f(b,a), f(j,m), f(a,i)), (f(i,b), f(j,m), f(a,i)): an outer pair times
    // any middle bucket.
    std::vector<char> mid(static_cast<std::size_t>(nf), 0), outer(static_cast<std::size_t>(nf_ * nf_), 0);
    for (int x = 0; x < N; ++x)
        for (int y = 0; y < N; ++y) mid[static_cast<std::size_t>(f(x, y))] = 1;
    // With (x, y, z) = (a, n, i) the first two patterns read the outer pairs
    // (f(x,y), f(y,z)) and (f(z,x), f(y,z)); with (x, y, z) = (b, a, i) the
    // last two read the same two forms.
    for (int x = 0; x < N; ++x)
        for (int y = 0; y < N; ++y)
            for (int z = 0; z < N; ++z) {
                outer[static_cast<std::size_t>(f(x, y) * nf_ + f(y, z))] = 1;
                outer[static_cast<std::size_t>(f(z, x) * nf_ + f(y, z))] = 1;
            }
    for (long o1 = 0; o1 < nf_; ++o1)
        for (long o3 = 0; o3 < nf_; ++o3)
            if (outer[static_cast<std::size_t>(o1 * nf_ + o3)])
                for (long m = 0; m < nf_; ++m)
                    if (mid[static_cast<std::size_t>(m)]) mark(o1, m, o3, 6);
    return need;
}

static std::size_t env_bytes(const char* name, std::size_t dflt) {
    const char* v = std::getenv(name);
    return v ? static_cast<std::size_t>(std::strtoull(v, nullptr, 10)) : dflt;
}

// One F/C/R pass made ready to launch: the job lists live in device memory,
// the launch geometry on the host.  Workspace pointers are taken from the
// context at launch time, so passes prepared one after another can share it.
struct FcrJob {
    FcrArgs P{};
    std::vector<int> cls_start;  // host copy of the Omega-class piece starts (A-job indices)
    std::vector<int> ajs;        // host copy of the A-job starts (sorted positions)
    std::vector<long> chunk_a;   // FFT chunks: A-jobs [chunk_a[c], chunk_a[c + 1])
    // combine batches: positions [comb0, comb0 + cb) holding chunks [c0, c1)
    struct Batch { long comb0, cb, c0, c1; };
    std::vector<Batch> batches;
    std::vector<long> paths;     // [ncomb + 1] prefix count of the F and C paths (traffic accounting)
    std::vector<long> outs;      // [ncomb + 1] prefix count of the F, C and R outputs
    std::vector<int> order;      // host copy of the combine list (position -> triple)
    std::vector<unsigned char> mask;
    long nfft_jobs = 0, ncomb = 0, chunk = 1, cconv = 1, spec_jobs = 0;
    std::size_t n = 0, Nt = 0, T = 0;
    int L = 0;
};

// Where a pass keeps its device job lists.
struct FcrLists {
    DevBuf* jobtri;
    DevBuf* jobmask;
    DevBuf* clsstart;
    DevBuf* jobrow;   // per sorted position: phase-table row (TACO_PHASE_LIBM)
    PhaseTab* phtab;  // the pass's phase table (TACO_PHASE_LIBM)
};

// TACO_PHASE_LIBM: the platform libm's exp(+i Omega (t dt)) for every
// distinct Omega of the pass (in the kernels' expression (w_i + w_j) + w_k),
// t < Nt, evaluated on the host and kept in the owner's table; rebuilt only
// when the Omegas, Nt or dt change.  Each sorted position gets its row.
struct PhaseRows {
    const double2* tab = nullptr;
    const int* rows = nullptr;
    long len = 0;
};

static std::size_t libm_table_budget() {
    const char* v = std::getenv("TACO_LIBM_TABLE_BYTES");
    return v ? static_cast<std::size_t>(std::strtoull(v, nullptr, 10)) : (std::size_t{16} << 30);
}

static PhaseRows phase_rows(taco_ctx* c, const FcrPlan& pl, const std::vector<int>& order,
                            const std::vector<std::uint64_t>& key, const FcrLists& lists) {
    PhaseRows out;
    if (phase_mode() != TACO_PHASE_LIBM || order.empty()) return out;
    std::vector<std::uint64_t> uniq(order.size());
    for (std::size_t q = 0; q < order.size(); ++q) uniq[q] = key[static_cast<std::size_t>(order[q])];
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<int> row(order.size());
    for (std::size_t q = 0; q < order.size(); ++q)
        row[q] = static_cast<int>(std::lower_bound(uniq.begin(), uniq.end(), key[static_cast<std::size_t>(order[q])]) -
                                  uniq.begin());
    const std::size_t len = pl.Nt;
    if (uniq.size() > libm_table_budget() / 16 / len)
        throw Error(TACO_ERR_INVALID, "libm phase table of " + std::to_string(uniq.size()) + " x " + std::to_string(len) +
                                          " values exceeds TACO_LIBM_TABLE_BYTES; use the TACO_PHASE_CR phase mode");
    std::uint64_t dt_bits;
    std::memcpy(&dt_bits, &pl.dt, 8);
    PhaseTab& T = *lists.phtab;
    if (!(T.valid && T.len == len && T.dt_bits == dt_bits && T.rows == uniq)) {
        T.valid = false;
        std::vector<double> om(uniq.size());
        for (std::size_t r = 0; r < uniq.size(); ++r) std::memcpy(&om[r], &uniq[r], 8);
        std::vector<double> host(2 * uniq.size() * len);
        libm_phase_table(om.data(), uniq.size(), len, pl.dt, host.data());
        T.tab.ensure(host.size() * sizeof(double));
        HIP_OK(hipMemcpyAsync(T.tab.p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));  // the host vector goes out of scope
        T.rows = std::move(uniq);
        T.len = len;
        T.dt_bits = dt_bits;
        T.valid = true;
    }
    lists.jobrow->ensure(row.size() * sizeof(int));
    HIP_OK(hipMemcpyAsync(lists.jobrow->p, row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    out.tab = T.tab.as<double2>();
    out.rows = lists.jobrow->as<int>();
    out.len = static_cast<long>(len);
    return out;
}

// A combine batch of cb triples runs split over its 256-sample blocks (three
// launches) when the triples alone would leave most of the chip idle; the
// block totals and carries take cb * nblk * 8 doubles in the context's
// `comb` buffer.
// TACO_COMBINE_SPLIT=0 / 1 forces the single-launch / split form where it fits.
static long combine_split_blocks(long cb, std::size_t T) {
    const long nblk = static_cast<long>((T + kBlock - 1) / kBlock);
    if (nblk < 2 || cb * nblk > (1L << 20)) return 0;
    const int force = env_int("TACO_COMBINE_SPLIT", -1);
    if (force >= 0) return force ? nblk : 0;
    return cb < 512 ? nblk : 0;
}

static FcrJob fcr_prepare(taco_ctx* c, const FcrPlan& pl, const FcrLists& lists) {
    // window mode: only the rows t < T enter the convolutions (the same causal
    // sums at t < T), with exact twiddles
    const bool window = fft_mode() == TACO_FFT_WINDOW;
    const std::size_t Nin = window ? std::max<std::size_t>(std::max(pl.T, pl.T_fft), 1) : pl.Nt;
    const std::size_t n = fcr_nfft(Nin, g_pad_factor);
    int L = 0;
    while ((std::size_t{1} << L) < n) ++L;
    if (L > 22) throw Error(TACO_ERR_INVALID, "FFT length above 2^22 (Nt above 2^21) is not supported");
    const int a = (L + 1) / 2, b = L - a;
    const Twiddles& tw = twiddles_for(c, L, window);
    const int nf = pl.nf;
    const long nf3 = static_cast<long>(nf) * nf * nf;

    // Triples to compute: every one (compute_triple_kernels) or the ones the
    // L4 assembly reads.  Those needing F or C get FFTs and are grouped by the
    // bit pattern of Omega = (w_i + w_j) + w_k (the kernels' expression) so one
    // class shares its phases exp(-i Omega t_k); R-only triples follow.
    std::vector<int> order;
    std::vector<unsigned char> mask;
    std::vector<std::uint64_t> key(static_cast<std::size_t>(nf3));
    triple_omega_keys(pl.h_omegas, nf, key.data());
    for (long t = 0; t < nf3; ++t) {
        const unsigned char m = pl.need ? (*pl.need)[static_cast<std::size_t>(t)] : 7;
        if (m & 3) order.push_back(static_cast<int>(t));
    }
    stable_sort_by_key(order, key.data());
    const long nfft_jobs = static_cast<long>(order.size());
    for (long t = 0; t < nf3; ++t) {
        const unsigned char m = pl.need ? (*pl.need)[static_cast<std::size_t>(t)] : 7;
        if (m == 4) order.push_back(static_cast<int>(t));
    }
    const long ncomb = static_cast<long>(order.size());
    mask.resize(order.size());
    for (std::size_t q = 0; q < order.size(); ++q)
        mask[q] = pl.need ? (*pl.need)[static_cast<std::size_t>(order[q])] : 7;
    // A-jobs: within a class the positions are in ascending triple order, so
    // the triples with the same i (the same phase-A input g_i exp(-i Omega t))
    // are consecutive; each run is one A-job.
    std::vector<int> ajs;
    for (long p = 0; p < nfft_jobs; ++p) {
        const int i = order[p] / (nf * nf);
        if (p == 0 || key[order[p]] != key[order[p - 1]] || i != order[p - 1] / (nf * nf))
            ajs.push_back(static_cast<int>(p));
    }
    const long nA = static_cast<long>(ajs.size());
    ajs.push_back(static_cast<int>(nfft_jobs));
    // Class pieces over A-jobs; a class with more than TACO_CLASS_MAX A-jobs
    // (default 4) is cut into pieces, each recomputing the same phases, so no
    // phase-A workgroup loops much longer than the others.
    const int cmax = std::max(1, env_int("TACO_CLASS_MAX", 4));
    std::vector<int> cls_start;
    for (long a = 0, run = 0; a < nA; ++a, ++run) {
        if (a == 0 || key[order[ajs[a]]] != key[order[ajs[a - 1]]]) run = 0;
        if (run % cmax == 0) cls_start.push_back(static_cast<int>(a));
    }
    cls_start.push_back(static_cast<int>(nA));
    // Phase B visits the FFT jobs of each chunk grouped by j, so the triples
    // running at the same time read the same two spectra (of g_mirror(j) and
    // conj(g_j)) and share them through L2.  The chunk length is fixed here
    // (chunks start at multiples of it).
    const std::size_t per_job_b = 3 * n * 16;
    long chunk_len = static_cast<long>(std::max<std::size_t>(1, chunk_budget_bytes() / per_job_b));
    chunk_len = std::max(1L, std::min(chunk_len, std::max(nfft_jobs, 1L)));
    // FFT chunks: whole A-jobs, at most chunk_len triples (a longer A-job alone)
    std::vector<long> chunk_a{0};
    for (long a = 0; a < nA;) {
        long a1 = a + 1;
        while (a1 < nA && ajs[a1 + 1] - ajs[chunk_a.back()] <= chunk_len) ++a1;
        chunk_a.push_back(a1);
        a = a1;
    }
    // Phase B visits the A-jobs of each chunk grouped by the j of their first
    // triple, so the workgroups running at one time mostly read the same
    // spectra (of g_mirror(j) and conj g_j), from L2.
    // (a counting sort by j: stable, linear in the A-jobs)
    std::vector<int> bperm(static_cast<std::size_t>(nA));
    {
        std::vector<int> cnt(static_cast<std::size_t>(nf) + 1);
        auto j_of = [&](long x) { return (order[ajs[x]] / nf) % nf; };
        for (std::size_t c = 0; c + 1 < chunk_a.size(); ++c) {
            const long c0 = chunk_a[c], c1 = chunk_a[c + 1];
            std::fill(cnt.begin(), cnt.end(), 0);
            for (long x = c0; x < c1; ++x) ++cnt[static_cast<std::size_t>(j_of(x)) + 1];
            for (int q = 0; q < nf; ++q) cnt[static_cast<std::size_t>(q) + 1] += cnt[static_cast<std::size_t>(q)];
            for (long x = c0; x < c1; ++x)
                bperm[static_cast<std::size_t>(c0 + cnt[static_cast<std::size_t>(j_of(x))]++)] = static_cast<int>(x);
        }
    }
    lists.jobtri->ensure(std::max<std::size_t>(order.size() + bperm.size() + ajs.size(), 1) * sizeof(int));
    lists.jobmask->ensure(std::max<std::size_t>(mask.size(), 1));
    lists.clsstart->ensure(cls_start.size() * sizeof(int));
    if (!order.empty()) {
        HIP_OK(hipMemcpyAsync(lists.jobtri->p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                              c->stream));
        HIP_OK(hipMemcpyAsync(lists.jobmask->p, mask.data(), mask.size(), hipMemcpyHostToDevice, c->stream));
        if (!bperm.empty())
            HIP_OK(hipMemcpyAsync(lists.jobtri->as<int>() + order.size(), bperm.data(), bperm.size() * sizeof(int),
                                  hipMemcpyHostToDevice, c->stream));
    }
    HIP_OK(hipMemcpyAsync(lists.jobtri->as<int>() + order.size() + bperm.size(), ajs.data(), ajs.size() * sizeof(int),
                          hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(lists.clsstart->p, cls_start.data(), cls_start.size() * sizeof(int),
                          hipMemcpyHostToDevice, c->stream));
    const PhaseRows prow = phase_rows(c, pl, order, key, lists);

    // FFT chunk (intermediates of at most chunk triples) and combine batches
    // (conv slots): consecutive chunks while they fit cconv slots, then the
    // R-only triples in batches of cconv.
    long chunk = 1;
    for (std::size_t c = 0; c + 1 < chunk_a.size(); ++c)
        chunk = std::max(chunk, static_cast<long>(ajs[chunk_a[c + 1]] - ajs[chunk_a[c]]));
    const std::size_t Tpad = std::max<std::size_t>(pl.T, 1);
    long cconv = static_cast<long>(env_bytes("TACO_CONV_BYTES", std::size_t{4} << 30) / (2 * Tpad * 16));
    cconv = std::max(cconv, chunk);
    cconv = std::max(1L, std::min(cconv, std::max(ncomb, chunk)));
    std::vector<FcrJob::Batch> batches;
    {
        std::size_t c = 0;
        const std::size_t nch = chunk_a.size() - 1;
        while (c < nch) {
            const long comb0 = ajs[chunk_a[c]];
            std::size_t c1 = c + 1;
            while (c1 < nch && ajs[chunk_a[c1 + 1]] - comb0 <= cconv) ++c1;
            batches.push_back({comb0, ajs[chunk_a[c1]] - comb0, static_cast<long>(c), static_cast<long>(c1)});
            c = c1;
        }
        for (long p = nfft_jobs; p < ncomb; p += cconv)
            batches.push_back({p, std::min(cconv, ncomb - p), static_cast<long>(nch), static_cast<long>(nch)});
    }
    const long spec_jobs = 2L * nf;
    c->spec.ensure(static_cast<std::size_t>(spec_jobs) * n * 16);
    c->inter1.ensure(static_cast<std::size_t>(std::max(chunk, spec_jobs)) * n * 16);
    c->inter2.ensure(static_cast<std::size_t>(2 * chunk) * n * 16);
    c->conv.ensure(static_cast<std::size_t>(2 * cconv) * Tpad * 16);
    for (const FcrJob::Batch& bt : batches)
        if (const long nblk = combine_split_blocks(bt.cb, pl.T))
            c->comb.ensure(static_cast<std::size_t>(bt.cb * nblk) * 8 * sizeof(double));
    FcrJob J;
    FcrArgs& P = J.P;
    P.gamma = pl.gamma;
    P.omegas = pl.omegas;
    P.mirror = pl.mirror;
    P.nf = nf;
    P.Nt = static_cast<long>(pl.Nt);
    P.Nin = static_cast<long>(Nin);
    P.dt = pl.dt;
    P.L = L;
    P.a = a;
    P.b = b;
    P.n = static_cast<long>(n);
    P.twStd = tw.std_.as<double2>();
    P.twB = tw.b_.as<double2>();
    P.twC = tw.c_.as<double2>();
    P.job_tri = lists.jobtri->as<int>();
    P.bperm = P.job_tri + order.size();
    P.ajob_start = P.bperm + bperm.size();
    P.job_mask = lists.jobmask->as<unsigned char>();
    P.nfft_jobs = nfft_jobs;
    P.cls_start = lists.clsstart->as<int>();
    P.cconv = cconv;
    P.T = static_cast<long>(pl.T);
    P.ph_tab = prow.tab;
    P.job_row = prow.rows;
    P.ph_len = prow.len;
    J.cls_start = std::move(cls_start);
    J.ajs = std::move(ajs);
    J.chunk_a = std::move(chunk_a);
    J.batches = std::move(batches);
    J.paths.assign(static_cast<std::size_t>(ncomb) + 1, 0);
    J.outs.assign(static_cast<std::size_t>(ncomb) + 1, 0);
    for (long p = 0; p < ncomb; ++p) {
        const int fc = p < nfft_jobs ? (mask[p] & 1) + ((mask[p] >> 1) & 1) : 0;
        J.paths[p + 1] = J.paths[p] + fc;
        J.outs[p + 1] = J.outs[p] + fc + ((mask[p] >> 2) & 1);
    }
    J.order = order;
    J.mask = mask;
    J.nfft_jobs = nfft_jobs;
    J.ncomb = ncomb;
    J.chunk = chunk;
    J.cconv = cconv;
    J.spec_jobs = spec_jobs;
    J.n = n;
    J.Nt = pl.Nt;
    J.T = pl.T;
    J.L = L;
    return J;
}

// Launches one prepared pass on the context stream.  The first pass of an
// FFT length picks the tile configurations (autotune synchronises the
// stream); later passes of that length only launch kernels.
static void fcr_launch(taco_ctx* c, FcrJob& J, const CombineArgs& out) {
    FcrArgs& P = J.P;
    P.inter1 = c->inter1.as<double2>();
    P.inter2 = c->inter2.as<double2>();
    P.spec = c->spec.as<double2>();
    P.conv = c->conv.as<double2>();
    const int L = J.L;
    const long spec_jobs = J.spec_jobs;
    const std::vector<int>& cls_start = J.cls_start;
    // Tile configuration per phase: TACO_FFT_TILE=<0..6> forces one; otherwise
    // the first chunk of each FFT length is run with every configuration
    // (identical butterflies, identical results) and the fastest is kept.
    auto it = c->fft_choice.find(L);
    if (it == c->fft_choice.end()) it = c->fft_choice.emplace(L, std::array<int, 3>{-1, -1, -1}).first;
    auto& choice = it->second;
    const int forced = env_int("TACO_FFT_TILE", -1);
    if (forced >= 0 && forced < kNumFftCfg) choice = {forced, forced, forced};
    const double cn = static_cast<double>(J.n) * 16.0;      // one complex series of length Nfft
    const double cNt = static_cast<double>(J.Nt) * 16.0;
    const double cT = static_cast<double>(J.T) * 16.0;
    // Spectra of every Gamma column and its conjugate.
    P.njobs = static_cast<int>(spec_jobs);
    timed(c, K_SPECTRA, spec_jobs * (cNt + 3 * cn), [&] {
        launch_phase(choice[0] >= 0 ? choice[0] : 0, PH_A_SPEC, P, spec_jobs, c->stream);
        launch_phase(choice[1] >= 0 ? choice[1] : 0, PH_B_SPEC, P, spec_jobs, c->stream);
    });
    HIP_OK(hipGetLastError());
    auto class_of = [&](long aj) {
        return static_cast<int>(std::upper_bound(cls_start.begin(), cls_start.end(), static_cast<int>(aj)) -
                                cls_start.begin()) - 1;
    };
    for (const FcrJob::Batch& bt : J.batches) {
        const long comb0 = bt.comb0, cb = bt.cb;
        P.comb0 = comb0;
        for (long ci = bt.c0; ci < bt.c1; ++ci) {
            const long a0 = J.chunk_a[static_cast<std::size_t>(ci)], a1 = J.chunk_a[static_cast<std::size_t>(ci) + 1];
            const long job0 = J.ajs[static_cast<std::size_t>(a0)], nj = J.ajs[static_cast<std::size_t>(a1)] - job0;
            P.job0 = job0;
            P.njobs = static_cast<int>(nj);
            P.a0 = a0;
            P.na = static_cast<int>(a1 - a0);
            P.cls0 = class_of(a0);
            const long ncls = class_of(a1 - 1) - P.cls0 + 1;
            const double na = static_cast<double>(a1 - a0);
            if (choice[0] < 0 || choice[1] < 0 || choice[2] < 0) autotune(c, P, ncls, a1 - a0, nj, choice);
            // phase A: read the Gamma column, write the half-transformed signal, once per A-job
            timed(c, K_PHASE_A, na * (cNt + cn), [&] { launch_phase(choice[0], PH_A_TRIPLE, P, ncls, c->stream); });
            // F and C paths of the chunk (one or both per triple)
            const double np_ = static_cast<double>(J.paths[job0 + nj] - J.paths[job0]);
            // phase B: read each A-job's signal once and one spectrum per path, write one
            // half-inverse product per path
            timed(c, K_PHASE_B, na * cn + np_ * 2 * cn,
                  [&] { launch_phase(choice[1], PH_B_TRIPLE, P, a1 - a0, c->stream); });
            // phase C: read each path's product, write its convolution for t < T
            timed(c, K_PHASE_C, np_ * (cn + cT), [&] { launch_phase(choice[2], PH_C, P, nj, c->stream); });
            HIP_OK(hipGetLastError());
        }
        // combine: read three Gamma columns and each path's convolution over
        // t < T, write the needed F/C/R values at the kept times
        const double kept = out.out_mode == 0 ? static_cast<double>(out.npos) : static_cast<double>(J.T);
        const double cpaths = static_cast<double>(J.paths[comb0 + cb] - J.paths[comb0]);
        const double couts = static_cast<double>(J.outs[comb0 + cb] - J.outs[comb0]);
        const long nblk = combine_split_blocks(cb, J.T);
        timed(c, K_COMBINE, cb * 3 * cT + cpaths * cT + couts * 16.0 * kept, [&] {
            if (!nblk) {
                hipLaunchKernelGGL(k_combine, dim3(static_cast<unsigned>(cb)), dim3(kBlock), 0, c->stream, P, out);
                return;
            }
            double* tc = c->comb.as<double>();
            const dim3 grid(static_cast<unsigned>(cb), static_cast<unsigned>(nblk));
            hipLaunchKernelGGL(k_combine_tot, grid, dim3(kBlock), 0, c->stream, P, out, tc, static_cast<int>(nblk));
            hipLaunchKernelGGL(k_combine_carry, dim3(static_cast<unsigned>((cb * 8 + 63) / 64)), dim3(64), 0, c->stream,
                               tc, cb, static_cast<int>(nblk));
            hipLaunchKernelGGL(k_combine_blocks, grid, dim3(kBlock), 0, c->stream, P, out,
                               static_cast<const double*>(tc), static_cast<int>(nblk));
        });
        HIP_OK(hipGetLastError());
    }
}

// The Direct method over the pass's triples (every one, or the ones the L4
// assembly reads), phases from the libm table or correctly rounded.
static void fcr_direct(taco_ctx* c, const FcrPlan& pl, const CombineArgs& out) {
    const int nf = pl.nf;
    const long nf3 = static_cast<long>(nf) * nf * nf;
    std::vector<int> order;
    std::vector<unsigned char> mask;
    std::vector<std::uint64_t> key(static_cast<std::size_t>(nf3));
    for (long t = 0; t < nf3; ++t) {
        const int k = static_cast<int>(t % nf), j = static_cast<int>((t / nf) % nf), i = static_cast<int>(t / (nf * nf));
        const double om = (pl.h_omegas[i] + pl.h_omegas[j]) + pl.h_omegas[k];
        std::memcpy(&key[static_cast<std::size_t>(t)], &om, 8);
        const unsigned char m = pl.need ? (*pl.need)[static_cast<std::size_t>(t)] : 7;
        if (m) {
            order.push_back(static_cast<int>(t));
            mask.push_back(m);
        }
    }
    if (order.empty() || pl.T == 0) return;
    const PhaseRows prow = phase_rows(c, pl, order, key, FcrLists{&c->jobtri, &c->jobmask, &c->clsstart, &c->jobrow,
                                                                 &c->phtab});
    c->jobtri.ensure(order.size() * sizeof(int));
    c->jobmask.ensure(mask.size());
    HIP_OK(hipMemcpyAsync(c->jobtri.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->jobmask.p, mask.data(), mask.size(), hipMemcpyHostToDevice, c->stream));
    DirectArgs D{};
    D.gamma = pl.gamma;
    D.Nt = static_cast<long>(pl.Nt);
    D.nf = nf;
    D.omegas = pl.omegas;
    D.mirror = pl.mirror;
    D.dt = pl.dt;
    D.T = static_cast<long>(pl.T);
    D.tri = c->jobtri.as<int>();
    D.mask = c->jobmask.as<unsigned char>();
    D.ph_tab = prow.tab;
    D.row = prow.rows;
    D.ph_len = prow.len;
    const double T = static_cast<double>(pl.T);
    timed(c, K_COMBINE, static_cast<double>(order.size()) * 3.0 * T * 16.0, [&] {
        hipLaunchKernelGGL(k_fcr_direct, dim3(static_cast<unsigned>(order.size()),
                                              static_cast<unsigned>((pl.T + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, c->stream, D, out);
    });
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));  // the host lists go out of scope
}

static void run_fcr(taco_ctx* c, const FcrPlan& pl, const CombineArgs& out) {
    if (pl.method == TACO_FCR_DIRECT) {
        fcr_direct(c, pl, out);
        return;
    }
    FcrJob J = fcr_prepare(c, pl, FcrLists{&c->jobtri, &c->jobmask, &c->clsstart, &c->jobrow, &c->phtab});
    fcr_launch(c, J, out);
}

static std::size_t l4_lds_bytes(int N, int num_ops) {
    const std::size_t N2 = static_cast<std::size_t>(N) * N;
    return N2 * N2 * 16 + static_cast<std::size_t>(num_ops) * N2 * 16 + N2 * 4;
}

// Raises the fused L4 kernel's dynamic LDS limit when a launch needs more than
// 64 KiB; once per context and size, so a captured graph never calls it.
static void set_l4_lds(taco_ctx* c, std::size_t lds) {
    if (lds > 64 * 1024 && lds > c->l4_lds_set) {
        HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_l4), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(lds)));
        c->l4_lds_set = lds;
    }
}

// Scratch of the factorised L4: complex values per position, and positions
// per launch (at most 1 GiB of scratch per launch).
static long l4f_per_pos(int N, int nf, int num_ops) {
    const long N2 = static_cast<long>(N) * N, P = static_cast<long>(num_ops) * num_ops;
    return P * (2 * N2 + 2 * N2 * nf) + N2 * nf + N2;
}
static long l4f_chunk(int N, int nf, int num_ops, long npos) {
    const long per = l4f_per_pos(N, nf, num_ops);
    return std::max(1L, std::min(npos, static_cast<long>((std::size_t{1} << 30) / (per * 16))));
}

// Factorised L4 (any N), positions in launches of at most 1 GiB of scratch.
static void launch_l4f(taco_ctx* c, const L4Args& A, long npos) {
    const long N2 = static_cast<long>(A.N) * A.N;
    L4fArgs F{};
    F.p2f = A.p2f;
    F.ops = A.ops;
    F.N = A.N;
    F.nf = A.nf;
    F.num_ops = A.num_ops;
    F.nf3 = A.nf3;
    F.GW = nullptr;
    F.scr_per = l4f_per_pos(A.N, A.nf, A.num_ops);
    const long chunk = l4f_chunk(A.N, A.nf, A.num_ops, npos);
    c->misc.ensure(static_cast<std::size_t>(chunk * F.scr_per) * 16);  // sized at creation for batches
    F.scr = c->misc.as<double2>();
    const std::size_t lds = (static_cast<std::size_t>(A.num_ops) * N2) * 16 + static_cast<std::size_t>(N2) * 4;
    if (lds > 64 * 1024) throw Error(TACO_ERR_INVALID, "factorised L4 kernel: coupling operators exceed 64 KiB of LDS");
    const double N4 = static_cast<double>(N2) * N2;
    for (long p0 = 0; p0 < npos; p0 += chunk) {
        const long np = std::min(chunk, npos - p0);
        F.Fs = A.Fs + p0 * A.nf3;
        F.Cs = A.Cs + p0 * A.nf3;
        F.Rs = A.Rs + p0 * A.nf3;
        F.L4 = A.L4 + p0 * static_cast<long>(N4);
        if (A.GW) F.GW = A.GW + p0 * static_cast<long>(N4);
        timed(c, K_L4, np * (3.0 * A.nf3 + N4) * 16.0, [&] {
            launch_k_l4f(static_cast<unsigned>(np), F.N, lds, c->stream, F);
        });
        HIP_OK(hipGetLastError());
    }
}

// The factorised kernel is the default (faster at every N measured: 2.5x at
// N = 4, 8x at N = 8); TACO_L4_KERNEL=lds selects the O(N^6) LDS kernel,
// which sums in the reference's order (N <= 8).
static void launch_l4(taco_ctx* c, const L4Args& A, long npos) {
    if (npos <= 0) return;
    const std::size_t lds = l4_lds_bytes(A.N, A.num_ops);
    const char* kind = std::getenv("TACO_L4_KERNEL");
    if (lds > 160 * 1024 || !kind || std::strcmp(kind, "lds") != 0) {
        launch_l4f(c, A, npos);
        return;
    }
    set_l4_lds(c, lds);
    const double N4 = static_cast<double>(A.N) * A.N * A.N * A.N;
    // per time: read the three nf^3 slices, write one N^4 block
    timed(c, K_L4, npos * (3.0 * A.nf3 + N4) * 16.0, [&] {
        hipLaunchKernelGGL(k_l4, dim3(static_cast<unsigned>(npos)), dim3(kBlock), lds, c->stream, A);
    });
    HIP_OK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Validation shared by the builders (tcl4_fused_cuda.cu:820-851, 1107-1216).
// ---------------------------------------------------------------------------
static void validate_system(std::size_t Nt, int nf, const int* mirror, const int* p2f, int N, int num_ops,
                            double dt, int method) {
    if (method != TACO_FCR_CONVOLUTION && method != TACO_FCR_DIRECT)
        throw Error(TACO_ERR_INVALID, "unknown FCRMethod");
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "dt must be > 0");
    if (nf <= 0 || Nt == 0) throw Error(TACO_ERR_INVALID, "empty gamma_series");
    if (N <= 0) throw Error(TACO_ERR_INVALID, "map.N must be > 0");
    if (num_ops <= 0) throw Error(TACO_ERR_INVALID, "coupling_ops must be non-empty");
    if (mirror)
        for (int b = 0; b < nf; ++b)
            if (mirror[b] < 0 || mirror[b] >= nf) throw Error(TACO_ERR_INVALID, "mirror index out of range");
    if (p2f)
        for (int e = 0; e < N * N; ++e) {
            if (p2f[e] < 0)
                throw Error(TACO_ERR_RUNTIME, "map.pair_to_freq contains -1 (missing frequency buckets)");
            if (p2f[e] >= nf) throw Error(TACO_ERR_INVALID, "pair_to_freq entry >= nf");
        }
}

struct TimePlan {
    std::vector<std::size_t> uniq;   // sorted unique times
    std::vector<int> t2pos;          // [T]
    std::vector<std::size_t> req2u;  // requested -> uniq position
    std::size_t T = 0;
};

static TimePlan plan_times(const std::size_t* tidx, std::size_t ntau, std::size_t Nt) {
    TimePlan tp;
    std::vector<std::size_t> req;
    if (ntau == 0) {
        req.resize(Nt);
        std::iota(req.begin(), req.end(), std::size_t{0});
    } else {
        req.assign(tidx, tidx + ntau);
    }
    for (std::size_t t : req)
        if (t >= Nt) throw Error(TACO_ERR_RANGE, "time_index out of range");
    tp.uniq = req;
    std::sort(tp.uniq.begin(), tp.uniq.end());
    tp.uniq.erase(std::unique(tp.uniq.begin(), tp.uniq.end()), tp.uniq.end());
    tp.T = tp.uniq.empty() ? 0 : tp.uniq.back() + 1;
    tp.t2pos.assign(tp.T, -1);
    for (std::size_t q = 0; q < tp.uniq.size(); ++q) tp.t2pos[tp.uniq[q]] = static_cast<int>(q);
    tp.req2u.resize(req.size());
    for (std::size_t q = 0; q < req.size(); ++q) tp.req2u[q] = static_cast<std::size_t>(tp.t2pos[req[q]]);
    return tp;
}

// Bytes of F/C/R slices one pass of the batch builders may hold:
// taco_set_slice_budget, else TACO_SLICE_BYTES, else 32 GiB.
static std::size_t slice_budget(const taco_ctx* c) {
    if (c->slice_budget) return c->slice_budget;
    static const std::size_t env = [] {
        const char* v = std::getenv("TACO_SLICE_BYTES");
        return v ? static_cast<std::size_t>(std::strtoull(v, nullptr, 10)) : std::size_t{0};
    }();
    return env ? env : (std::size_t{32} << 30);
}

// Core of the batch builders on device-resident inputs.  Writes L4 for the
// unique times into c->l4 and returns the plan; fcr_ms optional.
//
// The slices of one pass hold 3 nf^3 values per kept time.  When all kept
// times would exceed the slice budget, the kept times are cut into windows of
// consecutive positions and each window is one F/C/R pass over t < (its last
// time + 1) followed by the L4 of its times -- the time decomposition of the
// reference's distributed builder (tcl4_mpi_omp.cpp:62-156, each rank slicing
// its own times out of full series).  Every kept value is the same arithmetic
// as in one pass: the convolutions, prefix scans and phases at time t do not
// depend on how far the pass runs (window FFT mode keeps Nfft from the last
// kept time of the whole call), so windowed results are bit-identical.
static TimePlan l4_core(taco_ctx* c, const double2* d_gamma, std::size_t Nt, int nf, const double* d_omegas,
                        const double* h_omegas, const int* d_mirror, const int* d_p2f, const int* h_p2f, int N,
                        const double2* d_ops, int num_ops, double dt, const std::size_t* tidx, std::size_t ntau,
                        double* fcr_ms, int method = TACO_FCR_CONVOLUTION) {
    TimePlan tp = plan_times(tidx, ntau, Nt);
    const long npos = static_cast<long>(tp.uniq.size());
    if (npos == 0) return tp;
    const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
    const std::size_t N4 = static_cast<std::size_t>(N) * N * N * N;
    if (nf3 > 0 && static_cast<std::size_t>(npos) > (~std::size_t{0} / 48) / nf3)
        throw Error(TACO_ERR_OVERFLOW, "FCR buffer too large");
    const std::size_t per_pos = 3 * nf3 * 16;
    const long win = static_cast<long>(
        std::max<std::size_t>(1, std::min<std::size_t>(static_cast<std::size_t>(npos), slice_budget(c) / per_pos)));
    const long nwin = (npos + win - 1) / win;
    c->slices.ensure(per_pos * static_cast<std::size_t>(win));
    c->l4.ensure(N4 * static_cast<std::size_t>(npos) * 16);
    // t -> position within its window, one [T_w] block per window
    std::vector<int> t2w;
    std::vector<std::size_t> woff(static_cast<std::size_t>(nwin) + 1, 0);
    for (long w = 0; w < nwin; ++w) {
        const long p0 = w * win, p1 = std::min(npos, p0 + win);
        const std::size_t Tw = tp.uniq[static_cast<std::size_t>(p1 - 1)] + 1;
        woff[static_cast<std::size_t>(w)] = t2w.size();
        for (std::size_t t = 0; t < Tw; ++t) {
            const int q = tp.t2pos[t];
            t2w.push_back(q >= p0 && q < p1 ? static_cast<int>(q - p0) : -1);
        }
    }
    woff[static_cast<std::size_t>(nwin)] = t2w.size();
    c->t2pos.ensure(t2w.size() * sizeof(int));
    HIP_OK(hipMemcpyAsync(c->t2pos.p, t2w.data(), t2w.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const std::vector<unsigned char> need = needed_triples(h_p2f, N, nf);
    if (fcr_ms) *fcr_ms = 0.0;
    for (long w = 0; w < nwin; ++w) {
        const long p0 = w * win, np = std::min(npos, p0 + win) - p0;
        const std::size_t Tw = woff[static_cast<std::size_t>(w) + 1] - woff[static_cast<std::size_t>(w)];
        CombineArgs out{};
        out.out_mode = 0;
        out.t2pos = c->t2pos.as<int>() + woff[static_cast<std::size_t>(w)];
        out.npos = np;
        out.nf3 = static_cast<long>(nf3);
        double2* sl = c->slices.as<double2>();
        out.Fo = sl;
        out.Co = sl + nf3 * np;
        out.Ro = sl + 2 * nf3 * np;
        if (fcr_ms) HIP_OK(hipEventRecord(c->ev0, c->stream));
        if (env_int("TACO_DEBUG_POISON", 0))  // NaN-fill the slices: an entry L4 reads but nobody wrote shows up
            HIP_OK(hipMemsetAsync(c->slices.p, 0xff, per_pos * static_cast<std::size_t>(np), c->stream));
        FcrPlan pl{d_gamma, Nt, nf, d_omegas, h_omegas, d_mirror, dt, Tw, &need, tp.T, method};
        run_fcr(c, pl, out);
        if (fcr_ms) HIP_OK(hipEventRecord(c->ev1, c->stream));
        L4Args A{};
        A.Fs = out.Fo;
        A.Cs = out.Co;
        A.Rs = out.Ro;
        A.p2f = d_p2f;
        A.ops = d_ops;
        A.N = N;
        A.nf = nf;
        A.num_ops = num_ops;
        A.nf3 = static_cast<long>(nf3);
        A.L4 = c->l4.as<double2>() + static_cast<std::size_t>(p0) * N4;
        A.GW = nullptr;
        launch_l4(c, A, np);
        if (fcr_ms) {
            HIP_OK(hipEventSynchronize(c->ev1));
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            *fcr_ms += ms;
        }
    }
    return tp;
}

template <class T>
static T* upload(taco_ctx* c, DevBuf& buf, const T* host, std::size_t count) {
    buf.ensure(std::max<std::size_t>(count, 1) * sizeof(T));
    if (count) HIP_OK(hipMemcpyAsync(buf.p, host, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return buf.as<T>();
}

static void check_ctx(taco_ctx* c) {
    if (!c) throw Error(TACO_ERR_INVALID, "null taco_ctx");
}

// ---------------------------------------------------------------------------
// Batched run: many independent systems (uneven N, bath, grid length, time
// set) with every input, list and output resident in HBM.  Each system's
// F/C/R pass, fused L4 and TCL2 assembly are enqueued back to back with no
// host round trip, and one RK4 launch steps every system (a workgroup each).
// The whole sequence is captured once into a hipGraph and replayed.
// ---------------------------------------------------------------------------
struct BatchSys {
    std::size_t Nt = 0, T = 0, ntau = 0, steps = 0;
    int nf = 0, N = 0, num_ops = 0, order = 4;
    double dt = 0.0;
    long npos = 0;
    DevBuf gamma, omegas, mirror, p2f, ops, eps, r0;
    DevBuf jobtri, jobmask, clsstart, jobrow, t2pos, rows_out, rows_gen;
    PhaseTab phtab;  // this system's phase table (TACO_PHASE_LIBM)
    DevBuf l4out;  // [ntau][N^4] at the requested times, request order
    DevBuf lgen;   // [steps + 1][N^4] generator series of the RK4 run
    DevBuf rho;    // [steps + 1][N^2] samples, and the final vector after them
    FcrJob fcr;
};

struct BatchRun {
    std::vector<std::unique_ptr<BatchSys>> sys;
    DevBuf rkargs;
    long nrk = 0;
    std::size_t rk_lds = 0;
    int rk_block = 64;  // raised to the workgroup size each system needs
    double rk_bytes = 0.0;  // algorithmic bytes of the RK4 launch (generators read, samples written)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long gen = 0;  // g_dev_realloc when the graph was captured
    bool ran_eager = false;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~BatchRun() { drop_graph(); }
};

// Enqueues one system of a batch: F/C/R at its kept times, fused L4 into the
// context scratch, rows copied to the outputs, TCL2 added to the generator.
static void batch_enqueue_system(taco_ctx* c, BatchSys& S) {
    const std::size_t N2 = static_cast<std::size_t>(S.N) * S.N, N4 = N2 * N2;
    const std::size_t nf3 = static_cast<std::size_t>(S.nf) * S.nf * S.nf;
    if (S.npos > 0) {
        CombineArgs out{};
        out.out_mode = 0;
        out.t2pos = S.t2pos.as<int>();
        out.npos = S.npos;
        out.nf3 = static_cast<long>(nf3);
        double2* sl = c->slices.as<double2>();
        out.Fo = sl;
        out.Co = sl + nf3 * S.npos;
        out.Ro = sl + 2 * nf3 * S.npos;
        fcr_launch(c, S.fcr, out);
        L4Args A{};
        A.Fs = out.Fo;
        A.Cs = out.Co;
        A.Rs = out.Ro;
        A.p2f = S.p2f.as<int>();
        A.ops = S.ops.as<double2>();
        A.N = S.N;
        A.nf = S.nf;
        A.num_ops = S.num_ops;
        A.nf3 = static_cast<long>(nf3);
        A.L4 = c->l4.as<double2>();
        A.GW = nullptr;
        launch_l4(c, A, S.npos);
        if (S.ntau)
            hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>(S.ntau)), dim3(kBlock), 0, c->stream,
                               c->l4.as<double2>(), S.rows_out.as<int>(), static_cast<long>(N4),
                               S.l4out.as<double2>());
    }
    if (S.steps) {
        const long gen = static_cast<long>((S.steps + 1) * N4);
        if (S.order == 4)
            hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>(S.steps + 1)), dim3(kBlock), 0, c->stream,
                               c->l4.as<double2>(), S.rows_gen.as<int>(), static_cast<long>(N4), S.lgen.as<double2>());
        else
            hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((gen + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               c->stream, S.lgen.as<double2>(), gen);
        Tcl2Args A2{};
        A2.gamma = S.gamma.as<double2>();
        A2.Nt = static_cast<long>(S.Nt);
        A2.eps = S.eps.as<double>();
        A2.ops = S.ops.as<double2>();
        A2.p2f = S.p2f.as<int>();
        A2.N = S.N;
        A2.use_lists = tcl2_lists_enabled() ? 1 : 0;
        A2.nf = S.nf;
        A2.num_ops = S.num_ops;
        A2.L = S.lgen.as<double2>();
        A2.order = S.order;
        timed(c, K_TCL2, static_cast<double>(gen) * 32.0, [&] {
            launch_tcl2(static_cast<unsigned>(S.steps + 1), S.N, tcl2_lds(S.N), c->stream, A2);
        });
    }
    HIP_OK(hipGetLastError());
}

static void batch_enqueue(taco_ctx* c, BatchRun& B) {
    for (auto& s : B.sys) batch_enqueue_system(c, *s);
    if (B.nrk) {
        timed(c, K_RK4, B.rk_bytes, [&] {
            hipLaunchKernelGGL(k_rk4_batch, dim3(static_cast<unsigned>(B.nrk)), dim3(B.rk_block), B.rk_lds, c->stream,
                               B.rkargs.as<Rk4Sys>());
        });
        HIP_OK(hipGetLastError());
    }
}

}  // namespace taco_engine


// ---------------------------------------------------------------------------
// Forward tape and reverse pass (no reference counterpart: the reference has
// no derivatives).  The tape keeps on the device what the reverse pass reads:
// the F/C/R slices at every t <= steps, the generators L(t), and the RK4
// states every K steps (checkpoints; the states between are recomputed).
// ---------------------------------------------------------------------------
struct taco_tape {
    taco_ctx* ctx = nullptr;
    std::size_t Nt = 0, T = 0, steps = 0;
    int nf = 0, N = 0, num_ops = 0, order = 4, K = 64;
    double dt = 0.0;
    // Gamma made on the device from a bath (taco_tape_forward_bath): the
    // spectral density and its parameters, for the bath reverse
    bool bath = false;
    int bath_kind = 0;
    double bath_p0 = 0.0, bath_p1 = 0.0, bath_beta = 0.0;
    std::vector<double> h_eps, h_omegas;
    std::vector<int> h_p2f, h_mirror;
    std::vector<std::complex<double>> h_ops;
    taco_engine::TapeBufs bufs;
    taco_engine::DevBuf &gamma = bufs.gamma, &omegas = bufs.omegas, &mirror = bufs.mirror, &p2f = bufs.p2f,
                        &ops = bufs.ops, &eps = bufs.eps, &t2pos = bufs.t2pos, &jobtri = bufs.jobtri,
                        &jobmask = bufs.jobmask, &clsstart = bufs.clsstart, &slices = bufs.slices, &lgen = bufs.lgen,
                        &ck = bufs.ck;
    taco_engine::FcrJob fcr;
    bool fcr_ready = false;  // fcr prepared (job lists and phase table on the device)
};

namespace taco_engine {

// LDS of the RK4 workgroups: `vecs` vectors of D, then generators staged
// nstage at a time (column stride D + 1) within the 160 KiB of a CU.
constexpr std::size_t kRkLds = 160 * 1024;

static int rk4_nstage(int D, int vecs, long max_mats) {
    const std::size_t per = static_cast<std::size_t>(D) * (D + 1) * 16;
    const std::size_t base = static_cast<std::size_t>(vecs) * D * 16;
    if (base + 2 * per > kRkLds) return 0;
    long n = static_cast<long>((kRkLds - base) / per);
    n = std::min(n, std::min(max_mats, 256L));
    return n >= 2 ? static_cast<int>(n) : 0;
}

static std::size_t rk4_lds(int D, int vecs, int nstage) {
    return (static_cast<std::size_t>(vecs) * D + static_cast<std::size_t>(nstage) * D * (D + 1)) * 16;
}

// Threads of an RK4 workgroup: TACO_RK4_BLOCK (64..256, multiple of 64),
// else 256 (measured faster than one wave for D = 16..64).
static int env_int(const char* name, int dflt);
static int rk4_block(int D) {
    (void)D;
    const int b = env_int("TACO_RK4_BLOCK", 0);
    if (b >= 64 && b <= kBlock && b % 64 == 0) return b;
    return kBlock;
}

// Midpoint series of large-D generators kept while it fits 2 GiB.
static bool rk4_midpoints_fit(int D, std::size_t steps) {
    return steps * static_cast<std::size_t>(D) * D * 16 <= (std::size_t{2} << 30);
}
static const double2* rk4_midpoints(taco_ctx* c, DevBuf& buf, const double2* Ls, std::size_t steps, int D) {
    const long D2 = static_cast<long>(D) * D;
    buf.ensure(steps * static_cast<std::size_t>(D2) * 16);
    const long total = static_cast<long>(steps) * D2;
    const unsigned grid = static_cast<unsigned>(std::min<long>((total + kBlock - 1) / kBlock, 65536));
    hipLaunchKernelGGL(k_midpoints, dim3(grid), dim3(kBlock), 0, c->stream, Ls, static_cast<long>(steps), D2,
                       buf.as<double2>());
    HIP_OK(hipGetLastError());
    return buf.as<double2>();
}

static void rk4_launch(taco_ctx* ctx, const double2* dLs, const double2* dLh, std::size_t steps, int D, double2* dr,
                       double dt, std::size_t every, double2* dsamples, long* dns) {
    // algorithmic bytes: every generator read once, every sample written once
    const double gen_bytes = static_cast<double>(steps + 1) * D * D * 16.0 * (dLh ? 2.0 : 1.0);
    const double rk_bytes = gen_bytes + (dsamples ? static_cast<double>(steps / every + 2) * D * 16.0 : 0.0);
    if (!dLh && rk4_rows_threads(D) > 0) {
        timed(ctx, K_RK4, rk_bytes, [&] {
            rows_dispatch(D, [&](auto dc) {
                constexpr int DT = decltype(dc)::value;
                hipLaunchKernelGGL(k_rk4_rows<DT>, dim3(1), dim3(rk4_rows_threads(DT)), 3 * DT * 16, ctx->stream, dLs,
                                   static_cast<long>(steps), dr, dt, static_cast<long>(every), dsamples, dns);
            });
        });
        HIP_OK(hipGetLastError());
        return;
    }
    int nstage = dLh ? 0 : rk4_nstage(D, 6, static_cast<long>(steps) + 1);
    if (!dLh && nstage < 2 && steps > 0 && rk4_midpoints_fit(D, steps)) {
        // generators read from global memory: the midpoints once, so each
        // step reads four matrices instead of six (the same values)
        dLh = rk4_midpoints(ctx, ctx->rkmid, dLs, steps, D);
        nstage = 0;
    }
    const std::size_t lds = rk4_lds(D, 6, nstage);
    if (lds > kRkLds) throw Error(TACO_ERR_INVALID, "RK4 kernel supports D <= 1706");
    timed(ctx, K_RK4, rk_bytes, [&] {
        hipLaunchKernelGGL(k_rk4, dim3(1), dim3(rk4_block(D)), lds, ctx->stream, dLs, dLh, static_cast<long>(steps), D,
                           dr, dt, static_cast<long>(every), dsamples, dns, nstage);
    });
    HIP_OK(hipGetLastError());
}

static std::size_t l4_adj_lds(int N, int num_ops, std::size_t nf3, bool lds_slices, bool lds_tb = true) {
    const std::size_t N2 = static_cast<std::size_t>(N) * N;
    return ((lds_tb ? N2 * N2 : 0) + 2 * num_ops * N2 + (lds_slices ? 3 * nf3 : 0)) * 16 + N2 * 4;
}

static void set_fn_lds(const void* fn, std::size_t lds) {
    if (lds > 64 * 1024)
        HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
}

// F/C/R slices of every t < T into sl ([3][T][nf3]): the tape's pass,
// prepared on its first call and only relaunched after it.
static void tape_slices(taco_tape& tp, double2* sl) {
    taco_ctx* c = tp.ctx;
    const std::size_t T = tp.T;
    const std::size_t nf3 = static_cast<std::size_t>(tp.nf) * tp.nf * tp.nf;
    if (!tp.fcr_ready) {
        std::vector<int> t2pos(T);
        std::iota(t2pos.begin(), t2pos.end(), 0);
        tp.t2pos.ensure(T * sizeof(int));
        HIP_OK(hipMemcpy(tp.t2pos.p, t2pos.data(), T * sizeof(int), hipMemcpyHostToDevice));
        const std::vector<unsigned char> need = needed_triples(tp.h_p2f.data(), tp.N, tp.nf);
        FcrPlan pl{tp.gamma.as<double2>(), tp.Nt, tp.nf, tp.omegas.as<double>(), tp.h_omegas.data(),
                   tp.mirror.as<int>(), tp.dt, T, &need};
        tp.fcr = fcr_prepare(c, pl, FcrLists{&tp.jobtri, &tp.jobmask, &tp.clsstart, &tp.bufs.jobrow, &tp.bufs.phtab});
        tp.fcr_ready = true;
    }
    CombineArgs out{};
    out.out_mode = 0;
    out.t2pos = tp.t2pos.as<int>();
    out.npos = static_cast<long>(T);
    out.nf3 = static_cast<long>(nf3);
    out.Fo = sl;
    out.Co = sl + nf3 * T;
    out.Ro = sl + 2 * nf3 * T;
    fcr_launch(c, tp.fcr, out);
}

static std::size_t tape_slice_bytes(const taco_tape& tp) {
    const std::size_t nf3 = static_cast<std::size_t>(tp.nf) * tp.nf * tp.nf;
    if (tp.order != 4) return 0;
    if (nf3 > 0 && tp.T > (~std::size_t{0} / 48) / nf3) throw Error(TACO_ERR_OVERFLOW, "FCR buffer too large");
    return 3 * nf3 * tp.T * 16;
}

// Generators L(t) = L_TCL2(t) (+ L4(t)) of every t < T into tp.lgen, with
// the F/C/R slices written to sl.
static void tape_generator(taco_tape& tp, double2* sl) {
    taco_ctx* c = tp.ctx;
    const std::size_t N2 = static_cast<std::size_t>(tp.N) * tp.N, N4 = N2 * N2;
    const std::size_t T = tp.T;
    const std::size_t nf3 = static_cast<std::size_t>(tp.nf) * tp.nf * tp.nf;
    tp.lgen.ensure(T * N4 * 16);
    if (tp.order == 4) {
        tape_slices(tp, sl);
        L4Args A{};
        A.Fs = sl;
        A.Cs = sl + nf3 * T;
        A.Rs = sl + 2 * nf3 * T;
        A.p2f = tp.p2f.as<int>();
        A.ops = tp.ops.as<double2>();
        A.N = tp.N;
        A.nf = tp.nf;
        A.num_ops = tp.num_ops;
        A.nf3 = static_cast<long>(nf3);
        A.L4 = tp.lgen.as<double2>();
        A.GW = nullptr;
        launch_l4(c, A, static_cast<long>(T));
    } else {
        const long cnt = static_cast<long>(T * N4);
        hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, tp.lgen.as<double2>(), cnt);
    }
    Tcl2Args A2{};
    A2.gamma = tp.gamma.as<double2>();
    A2.Nt = static_cast<long>(tp.Nt);
    A2.eps = tp.eps.as<double>();
    A2.ops = tp.ops.as<double2>();
    A2.p2f = tp.p2f.as<int>();
    A2.N = tp.N;
    A2.use_lists = tcl2_lists_enabled() ? 1 : 0;
    A2.nf = tp.nf;
    A2.num_ops = tp.num_ops;
    A2.L = tp.lgen.as<double2>();
    A2.order = tp.order;
    launch_tcl2(static_cast<unsigned>(T), A2.N, tcl2_lds(A2.N), c->stream, A2);
    HIP_OK(hipGetLastError());
}

// Checkpoints r_{qK} of the sampled states ds ([T][D]) onto the tape.
static void tape_checkpoints(taco_tape& tp, const double2* ds) {
    taco_ctx* c = tp.ctx;
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N;
    const std::size_t nseg = (tp.steps + tp.K - 1) / tp.K;
    std::vector<int> rows(nseg);
    for (std::size_t q = 0; q < nseg; ++q) rows[q] = static_cast<int>(q * tp.K);
    tp.ck.ensure(std::max<std::size_t>(nseg, 1) * D * 16 + nseg * sizeof(int) + 16);
    int* drows = reinterpret_cast<int*>(tp.ck.as<double2>() + nseg * D);
    HIP_OK(hipMemcpyAsync(drows, rows.data(), nseg * sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>(nseg)), dim3(kBlock), 0, c->stream, ds, drows,
                       static_cast<long>(D), tp.ck.as<double2>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));  // the host row list goes out of scope
}

static void tape_forward(taco_tape& tp, const double* r0, double* samples_out) {
    taco_ctx* c = tp.ctx;
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N;
    const std::size_t T = tp.T;
    tp.slices.ensure(std::max<std::size_t>(tape_slice_bytes(tp), 16));
    tape_generator(tp, tp.slices.as<double2>());
    // RK4 with every state sampled; checkpoints r_{qK} kept on the tape.
    c->misc.ensure((T + 1) * D * 16 + 64);
    double2* dr = c->misc.as<double2>();
    double2* ds = dr + D;
    HIP_OK(hipMemcpyAsync(dr, r0, D * 16, hipMemcpyHostToDevice, c->stream));
    rk4_launch(c, tp.lgen.as<double2>(), nullptr, tp.steps, static_cast<int>(D), dr, tp.dt, 1, ds, nullptr);
    tape_checkpoints(tp, ds);
    if (samples_out) HIP_OK(hipMemcpyAsync(samples_out, ds, T * D * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
}

// Term lists of the factorised L4 reverse (gathers instead of scatter-adds):
// every X = C + R slice pair (x, z) = (f(b,a), f(a,i)) or (f(i,b), f(a,i))
// with its (a, i, b, form) terms; every F slice index with its M and I terms
// (n, i, a, b, kind) and every R slice index with its M and K terms, in a
// fixed order.  Packed 8 bits per index (N <= 128).
struct L4AdjLists {
    std::vector<int> xp_xz, xp_start, xp_ent, f_tgt, f_start, f_ent, r_tgt, r_start, r_ent;
};

// CSR of (key, word) pairs: the distinct keys ascending, each key's words in
// insertion order (a stable counting sort over keys 0..nkeys-1).
static void csr_by_key(const std::vector<int>& key, const std::vector<int>& word, long nkeys, std::vector<int>& keys,
                       std::vector<int>& start, std::vector<int>& ent) {
    std::vector<int> cnt(static_cast<std::size_t>(nkeys) + 1, 0);
    for (int k : key) ++cnt[static_cast<std::size_t>(k) + 1];
    for (long k = 0; k < nkeys; ++k) cnt[static_cast<std::size_t>(k) + 1] += cnt[static_cast<std::size_t>(k)];
    ent.assign(word.size(), 0);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (std::size_t q = 0; q < key.size(); ++q) ent[static_cast<std::size_t>(pos[static_cast<std::size_t>(key[q])]++)] = word[q];
    keys.clear();
    start.assign(1, 0);
    for (long k = 0; k < nkeys; ++k)
        if (cnt[static_cast<std::size_t>(k) + 1] > cnt[static_cast<std::size_t>(k)]) {
            keys.push_back(static_cast<int>(k));
            start.push_back(cnt[static_cast<std::size_t>(k) + 1]);
        }
}

static L4AdjLists l4_adjoint_lists(const int* p2f, int N, int nf) {
    if (N > 128) throw Error(TACO_ERR_INVALID, "factorised L4 reverse supports N <= 128");
    auto f = [&](int r, int c) { return p2f[r + c * N]; };
    const long nfl = nf;
    auto id3 = [nfl](long x, long y, long z) { return static_cast<int>((x * nfl + y) * nfl + z); };
    L4AdjLists L;
    // X-form pairs keyed z nf + x (the order of x + 65536 z), stored as x + 65536 z
    std::vector<int> xk, xw;
    xk.reserve(2 * static_cast<std::size_t>(N) * N * N);
    xw.reserve(xk.capacity());
    for (int a = 0; a < N; ++a)
        for (int i = 0; i < N; ++i)
            for (int b = 0; b < N; ++b) {
                const int w = a | (i << 8) | (b << 16);
                xk.push_back(f(a, i) * nf + f(b, a));
                xw.push_back(w);
                xk.push_back(f(a, i) * nf + f(i, b));
                xw.push_back(w | (1 << 24));
            }
    csr_by_key(xk, xw, nfl * nfl, L.xp_xz, L.xp_start, L.xp_ent);
    for (int& k : L.xp_xz) k = (k % nf) + 65536 * (k / nf);
    std::vector<int> fk, fw, rk, rw;
    const std::size_t n4 = static_cast<std::size_t>(N) * N * N * N;
    fk.reserve(2 * n4);
    fw.reserve(2 * n4);
    rk.reserve(2 * n4);
    rw.reserve(2 * n4);
    for (int n = 0; n < N; ++n)
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) {
                    const unsigned w = static_cast<unsigned>(n) | (static_cast<unsigned>(i) << 8) |
                                       (static_cast<unsigned>(a) << 16) | (static_cast<unsigned>(b) << 24);
                    const int w0 = static_cast<int>(w), w1 = static_cast<int>(w | (1u << 31));
                    fk.push_back(id3(f(b, a), f(b, i), f(n, b)));  // M: F part
                    fw.push_back(w0);
                    fk.push_back(id3(f(a, n), f(b, i), f(n, b)));  // I
                    fw.push_back(w1);
                    rk.push_back(id3(f(b, a), f(n, a), f(a, i)));  // M: R part
                    rw.push_back(w0);
                    rk.push_back(id3(f(i, b), f(n, a), f(a, i)));  // K
                    rw.push_back(w1);
                }
    csr_by_key(fk, fw, nfl * nfl * nfl, L.f_tgt, L.f_start, L.f_ent);
    csr_by_key(rk, rw, nfl * nfl * nfl, L.r_tgt, L.r_start, L.r_ent);
    return L;
}

// Reverse RK4 state of one tape: dLoss/dr_n (gb, [T][D]), the generator
// cotangents (Lb, [T][D^2]), the recomputed states and stage points of every
// step (seg, [steps][4][D]) and dLoss/dr_0 (rb, [D]).
struct RkRev {
    double2 *gb, *Lb, *seg, *rb;
    double2 *LsH = nullptr, *vec = nullptr;  // deferred walk: conjugate transposes, cotangent vectors
    double2* LhH = nullptr;                  // and their midpoints
};

// The walk reads transposed generators and leaves Lbar to k_rk4_lbar when it
// neither has a row-group kernel nor stages generators in LDS nor carries
// the L0 parts in registers (large D).
// TACO_RK4_DEFER=0 keeps the in-place walk (a test compares the two).
static bool rk_rev_deferred(const taco_tape& tp) {
    if (env_int("TACO_RK4_DEFER", 1) == 0) return false;
    const int D = tp.N * tp.N;
    if (D == 4 || D == 9 || D == 16) return false;
    if (static_cast<long>(D) * D <= static_cast<long>(kRkCarry) * kBlock) return false;
    return rk4_nstage(D, 12, static_cast<long>(tp.steps) + 1) < 2;
}

// The row-group walk (D = 4, 9, 16) stores kb1..kb4 per step and leaves
// Lbar to k_rk4_lbar as well (TACO_RK4_DEFER=0: Lbar formed in the walk).
static bool rk_rev_rows_deferred(const taco_tape& tp) {
    if (env_int("TACO_RK4_DEFER", 1) == 0) return false;
    const int D = tp.N * tp.N;
    return D == 4 || D == 9 || D == 16;
}

static std::size_t rk_rev_values(const taco_tape& tp) {
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N;
    const std::size_t base = tp.T * D + tp.T * D * D + 4 * tp.steps * D + D;
    if (rk_rev_rows_deferred(tp)) return base + 4 * tp.steps * D;
    return rk_rev_deferred(tp) ? base + tp.T * D * D + 4 * tp.steps * D + tp.steps * D * D : base;
}

static RkRev rk_rev_at(const taco_tape& tp, double2* base) {
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N;
    RkRev b{};
    b.gb = base;
    b.Lb = b.gb + tp.T * D;
    b.seg = b.Lb + tp.T * D * D;
    b.rb = b.seg + 4 * tp.steps * D;
    if (rk_rev_rows_deferred(tp)) {
        b.vec = b.rb + D;
    } else if (rk_rev_deferred(tp)) {
        b.LsH = b.rb + D;
        b.vec = b.LsH + tp.T * D * D;
        b.LhH = b.vec + 4 * tp.steps * D;
    }
    return b;
}

// Launch shape of the reverse RK4 kernels of a tape (the same for every
// tape of a D, steps and K): recompute (Rf) and walk (R) arguments, their
// LDS, and whether the walk keeps the next step's L0 part in registers
// (carry) instead of accumulating Lbar in place over a zeroed buffer.
struct RkRevShape {
    Rk4AdjArgs Rf, R;
    std::size_t lds_f = 0, lds_b = 0, nseg = 0;
    bool carry = false, rows_f = false, rows_b = false;
    double bytes = 0.0;
};

static RkRevShape rk_rev_shape(const taco_tape& tp, const RkRev& b) {
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N, D2 = D * D;
    const std::size_t K = static_cast<std::size_t>(tp.K), steps = tp.steps;
    RkRevShape sh;
    Rk4AdjArgs R{};
    R.Ls = tp.lgen.as<double2>();
    R.ck = tp.ck.as<double2>();
    R.gbar = b.gb;
    R.Lbar = b.Lb;
    R.seg = b.seg;
    R.r0bar = b.rb;
    R.steps = static_cast<long>(steps);
    R.D = static_cast<int>(D);
    R.K = tp.K;
    R.dt = tp.dt;
    R.LsH = b.LsH;
    R.vec = b.vec;
    R.LhH = b.LhH;
    sh.carry = D2 <= static_cast<std::size_t>(kRkCarry) * kBlock;
    sh.nseg = (steps + K - 1) / K;
    const int Di = static_cast<int>(D);
    sh.Rf = R;
    sh.Rf.nstage = rk4_nstage(Di, 6, static_cast<long>(K) + 1);
    sh.lds_f = rk4_lds(Di, 6, sh.Rf.nstage);
    sh.R = R;
    sh.R.nstage = rk4_nstage(Di, 12, static_cast<long>(steps) + 1);
    sh.lds_b = rk4_lds(Di, 12, sh.R.nstage);
    if (sh.lds_b > kRkLds || sh.lds_f > kRkLds) throw Error(TACO_ERR_INVALID, "adjoint RK4 kernel supports D <= 853");
    sh.rows_f = rk4_rows_threads(Di) > 0;
    sh.rows_b = Di == 4 || Di == 9 || Di == 16;
    // algorithmic bytes of the recompute and the walk: generators read twice,
    // stage points written and read, gbar read, Lbar written
    sh.bytes = 2.0 * static_cast<double>(steps + 1) * D2 * 16.0 + 8.0 * steps * D * 16.0 +
               static_cast<double>(tp.T) * (D + D2) * 16.0;
    return sh;
}

// Reverse RK4 of one tape: gb filled, Lb and rb written.
static void rk_rev_run(taco_tape& tp, const RkRev& b) {
    taco_ctx* c = tp.ctx;
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N, D2 = D * D;
    const RkRevShape sh = rk_rev_shape(tp, b);
    const int Di = static_cast<int>(D);
    if (b.LsH) {
        const int tiles = (Di + 15) / 16;
        hipLaunchKernelGGL(k_conj_transpose, dim3(static_cast<unsigned>(tiles * tiles), static_cast<unsigned>(tp.T)),
                           dim3(kBlock), 0, c->stream, tp.lgen.as<double2>(), b.LsH, Di, static_cast<long>(tp.T));
        const long D2l = static_cast<long>(D2), total = static_cast<long>(tp.steps) * D2l;
        hipLaunchKernelGGL(k_midpoints, dim3(static_cast<unsigned>(std::min<long>((total + kBlock - 1) / kBlock, 65536))),
                           dim3(kBlock), 0, c->stream, static_cast<const double2*>(b.LsH), static_cast<long>(tp.steps),
                           D2l, b.LhH);
    } else if (!sh.carry) {
        const long cntL = static_cast<long>(tp.T * D2);
        hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((cntL + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, b.Lb, cntL);
    }
    timed(c, K_RK4, sh.bytes, [&] {
        // every segment's states and stage points from its checkpoint, all segments at once
        if (sh.rows_f) {
            rows_dispatch(Di, [&](auto dc) {
                constexpr int DT = decltype(dc)::value;
                hipLaunchKernelGGL(k_rk4_recompute_rows<DT>, dim3(static_cast<unsigned>(sh.nseg)),
                                   dim3(rk4_rows_threads(DT)), 3 * DT * 16, c->stream, sh.Rf);
            });
        } else {
            hipLaunchKernelGGL(k_rk4_recompute, dim3(static_cast<unsigned>(sh.nseg)), dim3(kBlock), sh.lds_f,
                               c->stream, sh.Rf);
        }
        if (sh.rows_b) {
            rows_dispatch(Di, [&](auto dc) {
                constexpr int DT = decltype(dc)::value;
                if constexpr (DT <= 16)
                    hipLaunchKernelGGL(k_rk4_adjoint_rows<DT>, dim3(1), dim3(64), 5 * DT * 16, c->stream, sh.R);
            });
        } else {
            hipLaunchKernelGGL(k_rk4_adjoint, dim3(1), dim3(kBlock), sh.lds_b, c->stream, sh.R);
        }
        if (b.vec)
            hipLaunchKernelGGL(k_rk4_lbar, dim3(static_cast<unsigned>((D2 + kBlock - 1) / kBlock),
                                                static_cast<unsigned>(tp.T)),
                               dim3(kBlock), 0, c->stream, sh.R);
    });
    HIP_OK(hipGetLastError());
}

static void tape_adjoint_tail(taco_tape& tp, const RkRev& rv, const double2* sl, double* gamma_bar,
                              double* omega_bar, double* ops_bar, double* eps_bar, double* r0_bar,
                              double* bath_bar = nullptr);
static void tape_bath_adjoint(taco_tape& tp, const double2* gb4, const double2* g2, double* bath_bar,
                              std::vector<double>& omb);

// The term lists depend on (N, nf, pair_to_freq) only: built once per
// context and structure (the last 128 kept: a multi-set tape's parameter
// sets may each have their own bucket map), reused by every later reverse.
static const L4AdjLists& l4_adjoint_lists_cached(taco_ctx* c, const int* p2f, int N, int nf) {
    std::vector<int> key(p2f, p2f + static_cast<std::size_t>(N) * N);
    key.push_back(N);
    key.push_back(nf);
    for (auto& e : c->adj_lists)
        if (e.first == key) return *e.second;
    if (c->adj_lists.size() >= 128) c->adj_lists.erase(c->adj_lists.begin());
    c->adj_lists.emplace_back(std::move(key), std::make_shared<L4AdjLists>(l4_adjoint_lists(p2f, N, nf)));
    return *c->adj_lists.back().second;
}

static void tape_adjoint(taco_tape& tp, const double* samples_bar, double* gamma_bar, double* omega_bar,
                         double* ops_bar, double* eps_bar, double* r0_bar, double* bath_bar = nullptr) {
    taco_ctx* c = tp.ctx;
    const std::size_t D = static_cast<std::size_t>(tp.N) * tp.N;
    c->rk.ensure(rk_rev_values(tp) * 16);
    const RkRev rv = rk_rev_at(tp, c->rk.as<double2>());
    HIP_OK(hipMemcpyAsync(rv.gb, samples_bar, tp.T * D * 16, hipMemcpyHostToDevice, c->stream));
    rk_rev_run(tp, rv);
    tape_adjoint_tail(tp, rv, tp.slices.as<double2>(), gamma_bar, omega_bar, ops_bar, eps_bar, r0_bar, bath_bar);
}

// Everything after the reverse RK4: the L4 reverse over the slices sl, the
// F/C/R and TCL2 reverses, and the cotangents to the host (for a
// bath-sourced tape with bath_bar set, also the bath chain's reverse).
static void tape_adjoint_tail(taco_tape& tp, const RkRev& rv, const double2* sl, double* gamma_bar,
                              double* omega_bar, double* ops_bar, double* eps_bar, double* r0_bar,
                              double* bath_bar) {
    taco_ctx* c = tp.ctx;
    const std::size_t N2 = static_cast<std::size_t>(tp.N) * tp.N, D = N2, D2 = D * D;
    const std::size_t T = tp.T;
    const std::size_t nf3 = static_cast<std::size_t>(tp.nf) * tp.nf * tp.nf;
    double2* Lb = rv.Lb;
    double2* rb = rv.rb;
    // TCL2 reverse: per-(operator, bucket) parts and the device time sums
    const Tcl2Parts parts = tcl2_parts(tp.h_ops.data(), tp.num_ops, tp.N, tp.h_p2f.data(), tp.nf);
    const std::size_t npart = parts.al.size();
    const std::size_t nfz = static_cast<std::size_t>(tp.nf);
    const std::size_t nout = nfz * D2 + nfz * N2 + static_cast<std::size_t>(tp.N);
    const int nchunk = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(64, (std::size_t{1} << 22) / nout)));
    const std::size_t n_tcl2 = T * N2 + (T * tp.N + 1) / 2 + nfz * T + nchunk * nout + nout + npart * (D2 + 3 * N2) +
                               (npart + 1) / 4 + 1;
    c->t2scr.ensure(n_tcl2 * 16);
    double2* t2X = c->t2scr.as<double2>();
    double* t2eps = reinterpret_cast<double*>(t2X + T * N2);
    double2* t2g = t2X + T * N2 + (T * tp.N + 1) / 2;
    double2* t2part = t2g + nfz * T;
    double2* t2w = t2part + nchunk * nout;
    double2* t2K = t2w + nout;
    double2* t2A = t2K + npart * D2;
    double2* t2P = t2A + npart * N2;
    double2* t2Pb = t2P + npart * N2;
    int* t2b = reinterpret_cast<int*>(t2Pb + npart * N2);
    if (npart) {
        HIP_OK(hipMemcpyAsync(t2A, parts.AdA.data(), npart * N2 * 16, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(t2P, parts.P.data(), npart * N2 * 16, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(t2b, parts.b.data(), npart * sizeof(int), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_tcl2_part_k, dim3(static_cast<unsigned>((D2 + kBlock - 1) / kBlock),
                                               static_cast<unsigned>(npart)),
                           dim3(kBlock), 0, c->stream, t2P, t2A, tp.N, t2K);
        HIP_OK(hipGetLastError());
    }
    std::vector<std::complex<double>> gbar_h(static_cast<std::size_t>(tp.nf) * T, {0.0, 0.0});
    std::vector<std::complex<double>> opsb(static_cast<std::size_t>(tp.num_ops) * N2, {0.0, 0.0});
    std::vector<double> omb(static_cast<std::size_t>(tp.nf), 0.0), epsb(static_cast<std::size_t>(tp.N), 0.0);
    // The L4 part's host sums wait only for its copies (event l4_done), after
    // the TCL2 reverse is queued: the device runs that meanwhile.  Host
    // buffers of those copies live until then.
    std::vector<int> bl_start, bl_ent;        // bucket lists uploaded for k_bucket_reduce
    std::vector<double> omp;                  // per combine position omega partials
    std::vector<std::complex<double>> ap;     // per-time operator partials
    const std::vector<int>* comb_order = nullptr;
    hipEvent_t l4_done = nullptr;
    // on an error, the copies still queued finish before those buffers go
    struct SyncOnUnwind {
        hipStream_t s;
        hipEvent_t* e;
        ~SyncOnUnwind() {
            if (std::uncaught_exceptions() > 0) (void)hipStreamSynchronize(s);
            if (*e) (void)hipEventDestroy(*e);
        }
    } unwind_guard{c->stream, &l4_done};
    const double2* gb4_dev = nullptr;  // Gamma_bar of the L4 part on the device, when there is one
    if (bath_bar && !tp.bath) throw Error(TACO_ERR_INVALID, "bath cotangents need a bath-sourced tape");
    if (tp.order == 4) {
        const FcrJob& J = tp.fcr;
        const long ncomb = J.ncomb;
        // L4 adjoint: slice cotangents [T][3][nf3], operator partials [T][ops N^2]
        // factorised L4 adjoint (default; TACO_L4_ADJ_KERNEL=direct: the O(N^6) term walk)
        const bool l4_direct = [] {
            const char* v = std::getenv("TACO_L4_ADJ_KERNEL");
            return v && std::string(v) == "direct";
        }();
        // The N^4 Tbar tile stays in LDS while it fits (N <= 8); the factorised
        // kernel otherwise keeps it in its scratch (TACO_L4_ADJ_TB=global
        // forces that placement at any N).
        const bool tb_global_env = [] {
            const char* v = std::getenv("TACO_L4_ADJ_TB");
            return v && std::string(v) == "global";
        }();
        const bool lds_tb = l4_direct || (!tb_global_env && l4_adj_lds(tp.N, tp.num_ops, nf3, false) <= 160 * 1024);
        const bool lds_sl = l4_adj_lds(tp.N, tp.num_ops, nf3, true, lds_tb) <= 96 * 1024;
        const std::size_t lds = l4_adj_lds(tp.N, tp.num_ops, nf3, lds_sl, lds_tb);
        if (lds > 160 * 1024) throw Error(TACO_ERR_INVALID, "direct L4 adjoint kernel supports N <= 8");
        if (l4_direct) set_fn_lds(reinterpret_cast<const void*>(k_l4_adjoint), lds);
        const std::size_t sbar_n = T * 3 * nf3, abar_n = T * tp.num_ops * N2;
        const std::size_t nscr = static_cast<std::size_t>(ncomb) * 11 * T;
        const unsigned l4_grid = static_cast<unsigned>(std::min<std::size_t>(T, 1024));
        const std::size_t l4_per =
            4 * N2 * static_cast<std::size_t>(tp.nf) + 3 * N2 + N2 * N2 + (lds_tb ? 0 : N2 * N2);
        const std::size_t l4_scr = l4_direct ? 0 : l4_grid * l4_per;
        c->inter2.ensure((sbar_n + abar_n + nscr + static_cast<std::size_t>(tp.nf) * T + l4_scr) * 16 +
                         static_cast<std::size_t>(ncomb) * 8 + 64);
        double2* sbar = c->inter2.as<double2>();
        double2* abar = sbar + sbar_n;
        double2* scr = abar + abar_n;
        double2* gb4 = scr + nscr;
        double2* l4s = gb4 + static_cast<std::size_t>(tp.nf) * T;
        double* om = reinterpret_cast<double*>(l4s + l4_scr);
        if (!lds_sl)
            hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((sbar_n + kBlock - 1) / kBlock)), dim3(kBlock),
                               0, c->stream, sbar, static_cast<long>(sbar_n));
        L4AdjArgs A{};
        A.Lbar = Lb;
        A.Fs = sl;
        A.Cs = A.Fs + nf3 * T;
        A.Rs = A.Fs + 2 * nf3 * T;
        A.p2f = tp.p2f.as<int>();
        A.ops = tp.ops.as<double2>();
        A.N = tp.N;
        A.nf = tp.nf;
        A.num_ops = tp.num_ops;
        A.nf3 = static_cast<long>(nf3);
        A.sbar = sbar;
        A.Abar = abar;
        A.lds_slices = lds_sl ? 1 : 0;
        if (l4_direct) {
            timed(c, K_L4, 0.0, [&] {
                hipLaunchKernelGGL(k_l4_adjoint, dim3(static_cast<unsigned>(T)), dim3(kBlock), lds, c->stream, A);
            });
        } else {
            L4fAdjArgs Af{};
            Af.Lbar = A.Lbar;
            Af.Fs = A.Fs;
            Af.Cs = A.Cs;
            Af.Rs = A.Rs;
            Af.p2f = A.p2f;
            Af.ops = A.ops;
            Af.N = A.N;
            Af.nf = A.nf;
            Af.num_ops = A.num_ops;
            Af.nf3 = A.nf3;
            Af.T = static_cast<long>(T);
            Af.sbar = sbar;
            Af.Abar = abar;
            Af.lds_slices = A.lds_slices;
            Af.lds_tb = lds_tb ? 1 : 0;
            Af.scr = l4s;
            Af.scr_per = static_cast<long>(l4_per);
            const L4AdjLists& LL = l4_adjoint_lists_cached(c, tp.h_p2f.data(), tp.N, tp.nf);
            const std::vector<int>* parts_l[9] = {&LL.xp_xz, &LL.xp_start, &LL.xp_ent, &LL.f_tgt, &LL.f_start,
                                                  &LL.f_ent, &LL.r_tgt, &LL.r_start, &LL.r_ent};
            std::size_t tot = 0;
            for (auto* v : parts_l) tot += v->size();
            tp.bufs.adjl.ensure(std::max<std::size_t>(tot, 1) * sizeof(int));
            int* dl = tp.bufs.adjl.as<int>();
            const int* dp[9];
            for (int q = 0; q < 9; ++q) {
                dp[q] = dl;
                if (!parts_l[q]->empty())
                    HIP_OK(hipMemcpyAsync(dl, parts_l[q]->data(), parts_l[q]->size() * sizeof(int),
                                          hipMemcpyHostToDevice, c->stream));
                dl += parts_l[q]->size();
            }
            Af.xp_xz = dp[0];
            Af.xp_start = dp[1];
            Af.xp_ent = dp[2];
            Af.nxp = static_cast<int>(LL.xp_xz.size());
            Af.f_tgt = dp[3];
            Af.f_start = dp[4];
            Af.f_ent = dp[5];
            Af.nft = static_cast<int>(LL.f_tgt.size());
            Af.r_tgt = dp[6];
            Af.r_start = dp[7];
            Af.r_ent = dp[8];
            Af.nrt = static_cast<int>(LL.r_tgt.size());
            HIP_OK(hipStreamSynchronize(c->stream));  // the host lists go out of scope
            const bool large = tp.N >= kL4AdjLargeN;
            set_fn_lds(large ? reinterpret_cast<const void*>(k_l4f_adjoint<kL4AdjBlockLarge>)
                             : reinterpret_cast<const void*>(k_l4f_adjoint<kBlock>),
                       lds);
            timed(c, K_L4, 0.0, [&] {
                if (large)
                    hipLaunchKernelGGL(k_l4f_adjoint<kL4AdjBlockLarge>, dim3(l4_grid), dim3(kL4AdjBlockLarge), lds,
                                       c->stream, Af);
                else
                    hipLaunchKernelGGL(k_l4f_adjoint<kBlock>, dim3(l4_grid), dim3(kBlock), lds, c->stream, Af);
            });
        }
        HIP_OK(hipGetLastError());
        if (ncomb > 0) {
            FcrAdjArgs P{};
            P.gamma = tp.gamma.as<double2>();
            P.omegas = tp.omegas.as<double>();
            P.mirror = tp.mirror.as<int>();
            P.nf = tp.nf;
            P.Nt = static_cast<long>(tp.Nt);
            P.dt = tp.dt;
            P.T = static_cast<long>(T);
            P.job_tri = tp.jobtri.as<int>();
            P.job_mask = tp.jobmask.as<unsigned char>();
            P.ncomb = ncomb;
            P.sbar = sbar;
            P.nf3 = static_cast<long>(nf3);
            P.scr = scr;
            P.om = om;
            P.ph_tab = J.P.ph_tab;
            P.job_row = J.P.job_row;
            P.ph_len = J.P.ph_len;
            const unsigned nub = static_cast<unsigned>((T + kBlock - 1) / kBlock);
            timed(c, K_COMBINE, 0.0, [&] {
                hipLaunchKernelGGL(k_fcr_adj_scan, dim3(static_cast<unsigned>(ncomb)), dim3(kBlock), 0, c->stream, P);
                hipLaunchKernelGGL(k_fcr_adj_corr, dim3(static_cast<unsigned>(ncomb), nub), dim3(kBlock), 0,
                                   c->stream, P);
                hipLaunchKernelGGL(k_fcr_adj_fin, dim3(static_cast<unsigned>(ncomb)), dim3(kBlock), 0, c->stream, P);
            });
            HIP_OK(hipGetLastError());
            // bucket lists: (position, role) with role 0 g1 = Gamma_i, 1 g2 = Gamma_j, 2 gm = Gamma_mirror(j)
            std::vector<std::vector<int>> lists(static_cast<std::size_t>(tp.nf));
            for (long q = 0; q < ncomb; ++q) {
                const int tri = J.order[static_cast<std::size_t>(q)];
                const int i = tri / (tp.nf * tp.nf), j = (tri / tp.nf) % tp.nf;
                lists[static_cast<std::size_t>(i)].push_back(static_cast<int>(q * 3 + 0));
                lists[static_cast<std::size_t>(j)].push_back(static_cast<int>(q * 3 + 1));
                lists[static_cast<std::size_t>(tp.h_mirror[static_cast<std::size_t>(j)])].push_back(
                    static_cast<int>(q * 3 + 2));
            }
            std::vector<int>& start = bl_start;
            std::vector<int>& ent = bl_ent;
            start.assign(1, 0);
            for (auto& l : lists) {
                ent.insert(ent.end(), l.begin(), l.end());
                start.push_back(static_cast<int>(ent.size()));
            }
            c->misc.ensure((start.size() + ent.size()) * sizeof(int) + 16);
            int* dstart = c->misc.as<int>();
            int* dent = dstart + start.size();
            HIP_OK(hipMemcpyAsync(dstart, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
            HIP_OK(hipMemcpyAsync(dent, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
            // each bucket's list in kBredPieces contiguous pieces (partial sums in
            // list order), the pieces then summed in order
            c->bred.ensure(static_cast<std::size_t>(kBredPieces) * tp.nf * T * 16);
            hipLaunchKernelGGL(k_bucket_reduce, dim3(nub, static_cast<unsigned>(tp.nf), kBredPieces), dim3(kBlock), 0,
                               c->stream, scr, static_cast<long>(T), dstart, dent, c->bred.as<double2>());
            hipLaunchKernelGGL(k_bucket_reduce_pieces, dim3(nub, static_cast<unsigned>(tp.nf)), dim3(kBlock), 0,
                               c->stream, c->bred.as<double2>(), static_cast<long>(T), static_cast<long>(T), gb4);
            HIP_OK(hipGetLastError());
            gb4_dev = gb4;
            HIP_OK(hipMemcpyAsync(gbar_h.data(), gb4, gbar_h.size() * 16, hipMemcpyDeviceToHost, c->stream));
            omp.resize(static_cast<std::size_t>(ncomb));
            HIP_OK(hipMemcpyAsync(omp.data(), om, omp.size() * 8, hipMemcpyDeviceToHost, c->stream));
            comb_order = &J.order;
        }
        // operator cotangents of the L4 part: per-time partials, summed on the host below
        ap.resize(abar_n);
        HIP_OK(hipMemcpyAsync(ap.data(), abar, abar_n * 16, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipEventCreateWithFlags(&l4_done, hipEventDisableTiming));
        HIP_OK(hipEventRecord(l4_done, c->stream));
    }
    // TCL2 part: per-time reverse and time sums on the device, operator
    // cotangents on the host from the sums
    Tcl2AdjArgs Q{};
    Q.Lbar = Lb;
    Q.gamma = tp.gamma.as<double2>();
    Q.Nt = static_cast<long>(tp.Nt);
    Q.T = static_cast<long>(T);
    Q.K = t2K;
    Q.AdA = t2A;
    Q.part_b = t2b;
    Q.npart = static_cast<int>(npart);
    Q.N = tp.N;
    Q.nf = tp.nf;
    Q.order = tp.order;
    Q.gbar = t2g;
    Q.X = t2X;
    Q.epsd = t2eps;
    Q.nchunk = nchunk;
    Q.nout = static_cast<long>(nout);
    Q.part = t2part;
    Q.wsum = t2w;
    const unsigned nob = static_cast<unsigned>((nout + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_tcl2_adj_t, dim3(static_cast<unsigned>(T)), dim3(kBlock), 2 * N2 * 16, c->stream, Q);
    if (tp.order != 0 && npart) {
        // sum_e Re(conj(K_part) Lbar_t): tiles of 8 parts x 8 times over e-chunks of 16384
        const long echunk = 16384;
        const int nzc = static_cast<int>((D2 + echunk - 1) / echunk);
        c->bred.ensure(static_cast<std::size_t>(nzc) * npart * T * sizeof(double));
        double* kd = reinterpret_cast<double*>(c->bred.as<double2>());
        hipLaunchKernelGGL(k_tcl2_adj_kdot,
                           dim3(static_cast<unsigned>((npart + kKdTile - 1) / kKdTile),
                                static_cast<unsigned>((T + kKdTile - 1) / kKdTile), static_cast<unsigned>(nzc)),
                           dim3(kBlock), 0, c->stream, t2K, Lb, static_cast<long>(D2), static_cast<int>(npart),
                           static_cast<long>(T), echunk, kd);
        hipLaunchKernelGGL(k_tcl2_adj_gam,
                           dim3(static_cast<unsigned>((T + kBlock - 1) / kBlock), static_cast<unsigned>(tp.nf)),
                           dim3(kBlock), 0, c->stream, kd, static_cast<int>(npart), nzc, static_cast<long>(T), t2b,
                           t2g);
    }
    hipLaunchKernelGGL(k_tcl2_adj_w, dim3(nob, static_cast<unsigned>(nchunk)), dim3(kBlock), 0, c->stream, Q);
    hipLaunchKernelGGL(k_tcl2_adj_wsum, dim3(nob), dim3(kBlock), 0, c->stream, Q);
    HIP_OK(hipGetLastError());
    const std::size_t nW = nfz * D2, nV = nfz * N2;
    // operator cotangents of the TCL2 part, per part on the device
    if (tp.order != 0 && npart) {
        if (N2 > 1024) throw Error(TACO_ERR_INVALID, "TCL2 reverse supports N <= 32");
        hipLaunchKernelGGL(k_tcl2_ops_parts, dim3(static_cast<unsigned>(npart)), dim3(kBlock), 2 * N2 * 16, c->stream,
                           t2w, t2w + nW, t2P, t2b, tp.N, t2Pb);
        HIP_OK(hipGetLastError());
    }
    if (l4_done) {
        // the L4 part's host sums (omega partials in combine order, operator
        // partials in time order) while the TCL2 reverse runs
        const hipError_t we = hipEventSynchronize(l4_done);
        (void)hipEventDestroy(l4_done);
        l4_done = nullptr;
        HIP_OK(we);
        for (std::size_t q = 0; q < omp.size(); ++q) {
            const int tri = (*comb_order)[q];
            const int k = tri % tp.nf, j = (tri / tp.nf) % tp.nf, i = tri / (tp.nf * tp.nf);
            const double v = omp[q];
            omb[static_cast<std::size_t>(i)] += v;
            omb[static_cast<std::size_t>(j)] += v;
            omb[static_cast<std::size_t>(k)] += v;
        }
        for (std::size_t t = 0; t < T; ++t)
            for (std::size_t e = 0; e < opsb.size(); ++e) opsb[e] += ap[t * opsb.size() + e];
    }
    if (bath_bar) tape_bath_adjoint(tp, gb4_dev, tp.order != 0 ? t2g : nullptr, bath_bar, omb);
    std::vector<std::complex<double>> g2(nfz * T), wv(nV + static_cast<std::size_t>(tp.N)), r0b(D), pb(npart * N2);
    HIP_OK(hipMemcpyAsync(g2.data(), t2g, g2.size() * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(wv.data(), t2w + nW, wv.size() * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(r0b.data(), rb, D * 16, hipMemcpyDeviceToHost, c->stream));
    if (tp.order != 0 && npart)
        HIP_OK(hipMemcpyAsync(pb.data(), t2Pb, pb.size() * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int n = 0; n < tp.N; ++n) epsb[static_cast<std::size_t>(n)] += wv[nV + static_cast<std::size_t>(n)].real();
    if (tp.order != 0) {
        for (std::size_t e = 0; e < g2.size(); ++e) gbar_h[e] += g2[e];
        // each operator entry takes the Pbar of the part of its bucket (one part per entry)
        for (std::size_t q = 0; q < npart; ++q) {
            std::complex<double>* Ab = opsb.data() + static_cast<std::size_t>(parts.al[q]) * N2;
            for (std::size_t e = 0; e < N2; ++e)
                if (tp.h_p2f[e] == parts.b[q]) Ab[e] += pb[q * N2 + e];
        }
    }
    if (gamma_bar) {
        auto* g = reinterpret_cast<std::complex<double>*>(gamma_bar);
        std::fill(g, g + tp.Nt * static_cast<std::size_t>(tp.nf), std::complex<double>(0.0, 0.0));
        for (int b = 0; b < tp.nf; ++b)
            for (std::size_t t = 0; t < T; ++t)
                g[static_cast<std::size_t>(b) * tp.Nt + t] = gbar_h[static_cast<std::size_t>(b) * T + t];
    }
    if (omega_bar) std::copy(omb.begin(), omb.end(), omega_bar);
    if (ops_bar) std::copy(opsb.begin(), opsb.end(), reinterpret_cast<std::complex<double>*>(ops_bar));
    if (eps_bar) std::copy(epsb.begin(), epsb.end(), eps_bar);
    if (r0_bar) std::copy(r0b.begin(), r0b.end(), reinterpret_cast<std::complex<double>*>(r0_bar));
}

}  // namespace taco_engine

namespace taco_engine {
constexpr double kPiHost = 3.141592653589793238462643383279502884;

// Forward FFT of x [2^L values] on the device, the reference's radix-2
// recurrence FFT bit for bit: for L <= 2 in place on one thread (tw: room for
// 4 complex twiddles), else by the spectrum-mode phase kernels into c->spec.
// X[k] sits at ((k mod 2^a) << b) + (k >> a).
struct SpecOut {
    const double2* X;
    int a, b;
};
static SpecOut fft_forward_dev(taco_ctx* c, double2* x, int L, double2* tw_tiny) {
    const std::size_t n = std::size_t{1} << L;
    if (L <= 2) {
        std::vector<double> tab;
        make_twiddles(L, tab);
        HIP_OK(hipMemcpyAsync(tw_tiny, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_fft_tiny, dim3(1), dim3(64), 0, c->stream, x, L, tw_tiny);
        HIP_OK(hipStreamSynchronize(c->stream));  // the host table goes out of scope
        HIP_OK(hipGetLastError());
        return {x, L, 0};
    }
    const Twiddles& tw = twiddles_for(c, L);
    FcrArgs P{};
    P.gamma = x;
    P.Nt = static_cast<long>(n);
    P.Nin = static_cast<long>(n);
    P.L = L;
    P.a = (L + 1) / 2;
    P.b = L - P.a;
    P.n = static_cast<long>(n);
    P.twStd = tw.std_.as<double2>();
    P.twB = tw.b_.as<double2>();
    P.twC = tw.c_.as<double2>();
    c->inter1.ensure(n * 16);
    c->spec.ensure(n * 16);
    P.inter1 = c->inter1.as<double2>();
    P.spec = c->spec.as<double2>();
    P.njobs = 1;
    auto it = c->fft_choice.find(L);
    const int cfgA = it != c->fft_choice.end() && it->second[0] >= 0 ? it->second[0] : 2;
    const int cfgB = it != c->fft_choice.end() && it->second[1] >= 0 ? it->second[1] : 2;
    launch_phase(cfgA, PH_A_SPEC, P, 1, c->stream);
    launch_phase(cfgB, PH_B_SPEC, P, 1, c->stream);
    HIP_OK(hipGetLastError());
    return {P.spec, P.a, P.b};
}

// Bath correlation C(t_n), n <= N, into device memory dC (N + 1 values): the
// sampled KMS spectrum, its forward FFT by the three-pass kernels (the
// reference's radix-2 recurrence FFT) and the scale dw/pi.
static void bcf_device(taco_ctx* c, std::size_t N, double dt, int kind, double p0, double p1, double beta, double2* dC) {
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "bcf_fft: dt must be > 0");
    if (kind != 0 && kind != 1) throw Error(TACO_ERR_INVALID, "bcf_fft: unknown spectral density kind");
    const std::size_t Nt = N + 1;
    std::size_t Nfft = 2 * Nt;
    if ((Nfft & (Nfft - 1)) != 0) Nfft = next_pow2(Nfft);
    int L = 0;
    while ((std::size_t{1} << L) < Nfft) ++L;
    if (L > 22) throw Error(TACO_ERR_INVALID, "bcf_fft: FFT length above 2^22 (N above 2^21) is not supported");
    const double tf = dt * static_cast<double>(Nfft);
    const double domega = 2.0 * kPiHost / tf;
    const double omegaNy = kPiHost / dt;
    const bool zero_t = std::isinf(beta);
    const long half = static_cast<long>(Nfft / 2);
    const bool libm_ef = phase_mode() == TACO_PHASE_LIBM && kind == TACO_J_OHMIC;
    // misc: S [Nfft], twiddles of the tiny FFT [4], kms [half + 1], Ohmic factors [half]
    c->misc.ensure((Nfft + 4) * 16 + static_cast<std::size_t>(half + 1) * 8 * (libm_ef ? 2 : 1));
    BcfArgs B{};
    B.kind = kind;
    B.p0 = p0;
    B.p1 = p1;
    B.beta = beta;
    B.domega = domega;
    B.omegaNy = omegaNy;
    B.r = zero_t ? 0.0 : exp_in_mode(-beta * domega);
    B.eNy = zero_t ? 0.0 : exp_in_mode(beta * omegaNy);
    B.nfft = static_cast<long>(Nfft);
    B.S = c->misc.as<double2>();
    double2* d_tw = B.S + Nfft;
    double* kms = reinterpret_cast<double*>(d_tw + 4);
    B.kms = kms;
    if (libm_ef) {  // the Ohmic factor of J at every sampled frequency, by the host's libm
        // entries 1..max(posCount, 1): the DC term reads J(domega) even when posCount = 0
        std::vector<double> ef(static_cast<std::size_t>(half + 1), 0.0);
        for (long k = 1; k < std::max(half, 2L); ++k)
            ef[static_cast<std::size_t>(k)] = std::exp(-(domega * static_cast<double>(k)) / p1);
        double* d_ef = kms + half + 1;
        HIP_OK(hipMemcpyAsync(d_ef, ef.data(), ef.size() * 8, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        B.efac = d_ef;
        B.efac_ny = std::exp(-omegaNy / p1);
    }
    if (!zero_t) hipLaunchKernelGGL(k_bcf_kms, dim3(1), dim3(64), 0, c->stream, B.r, half - 1, kms);
    hipLaunchKernelGGL(k_bcf_spectrum, dim3(static_cast<unsigned>((Nfft + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, B);
    const SpecOut X = fft_forward_dev(c, B.S, L, d_tw);
    hipLaunchKernelGGL(k_bcf_gather, dim3(static_cast<unsigned>((Nt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, X.X, X.a, X.b, static_cast<long>(Nt), domega / kPiHost, dC);
    HIP_OK(hipGetLastError());
}

// Trapz Gamma [nf][Nc] from device C [Nc] and device bucket frequencies.
// d_steps: room for nf complex values; in TACO_PHASE_LIBM it receives the
// host's phase steps std::exp(i omega_b dt) (gamma.hpp:185).
static void trapz_device(taco_ctx* c, const double2* dC, std::size_t Nc, double dt, const double* h_om,
                         const double* d_om, int nf, double2* d_steps, double2* dG) {
    if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "gamma_trapz: dt must be > 0");
    if (Nc == 0 || nf <= 0) return;
    const double2* steps = nullptr;
    if (phase_mode() == TACO_PHASE_LIBM) {
        std::vector<std::complex<double>> st(static_cast<std::size_t>(nf));
        for (int b = 0; b < nf; ++b) st[static_cast<std::size_t>(b)] = expi_in_mode(h_om[b] * dt);
        HIP_OK(hipMemcpyAsync(d_steps, st.data(), st.size() * 16, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        steps = d_steps;
    }
    hipLaunchKernelGGL(k_gamma_trapz, dim3(static_cast<unsigned>(nf)), dim3(kBlock), 0, c->stream, dC,
                       static_cast<long>(Nc), dt, d_om, steps, dG);
    HIP_OK(hipGetLastError());
}
}  // namespace taco_engine

namespace taco_engine {

// Buffers of a destroyed tape go back to the context's pool (at most 256
// entries and 32 GiB) for the next tapes, so gradients of the same shape
// stop allocating after the first.
static void tape_pool_return(taco_tape& tp) {
    auto& pool = tp.ctx->tape_pool;
    std::size_t held = 0;
    for (auto& b : pool) held += b->bytes();
    const std::size_t add = tp.bufs.bytes();
    if (pool.size() >= 256 || held + add > (std::size_t{32} << 30)) return;
    pool.push_back(std::make_unique<TapeBufs>());
    pool.back()->swap(tp.bufs);
}

// A bath whose correlation and trapz Gamma the device makes for a tape.
struct BathSrc {
    int kind;
    double p0, p1, beta;
};

// Device layout of a bath-sourced tape's `bath` buffer, in complex slots:
// C [Nt], phase steps [nf], Gamma_bar [nf][T], Q [nf][T], the reverse FFT's
// input [nfft] and twiddles [4], omega_bar [nf], workgroup sums [3 nblk] and
// five results (doubles, two per slot).
struct BathLayout {
    std::size_t nfft = 0, nblk = 0;
    int L = 0;
    std::size_t steps = 0, gbar = 0, q = 0, x = 0, tw = 0, wbar = 0, part = 0, out = 0, total = 0;
};
static BathLayout bath_layout(std::size_t Nt, int nf, std::size_t T) {
    BathLayout B;
    B.nfft = 2 * Nt;
    if ((B.nfft & (B.nfft - 1)) != 0) B.nfft = next_pow2(B.nfft);
    while ((std::size_t{1} << B.L) < B.nfft) ++B.L;
    B.nblk = std::max<std::size_t>(1, (B.nfft / 2 - 1 + kBlock - 1) / kBlock);
    const std::size_t nfz = static_cast<std::size_t>(nf);
    B.steps = Nt;
    B.gbar = B.steps + nfz;
    B.q = B.gbar + nfz * T;
    B.x = B.q + nfz * T;
    B.tw = B.x + B.nfft;
    B.wbar = B.tw + 4;
    B.part = B.wbar + (nfz + 1) / 2;
    B.out = B.part + (3 * B.nblk + 1) / 2;
    B.total = B.out + 3;
    return B;
}

// Validates one tape's inputs and uploads them (the tape keeps host copies of
// what its reverse pass reads on the host).  With a bath, the Gamma table is
// made on the device instead (bcf_fft_fun and the trapz Gamma, bit-identical
// to taco_bath_gamma_hip) and in.gamma is not read.
static std::unique_ptr<taco_tape> tape_make(taco_ctx* ctx, const taco_tape_set& in, std::size_t checkpoint_every,
                                            const BathSrc* bath = nullptr) {
    const std::size_t Nt = in.Nt, steps = in.steps;
    const int nf = in.nf, N = in.N, num_ops = in.num_ops;
    if (!(in.order == 0 || in.order == 2 || in.order == 4)) throw Error(TACO_ERR_INVALID, "order must be 0, 2, or 4");
    validate_system(Nt, nf, in.mirror, in.pair_to_freq, N, num_ops, in.dt, TACO_FCR_CONVOLUTION);
    if ((!bath && !in.gamma) || !in.omegas || !in.mirror || !in.pair_to_freq || !in.ops || !in.eps || !in.r0)
        throw Error(TACO_ERR_INVALID, "tape inputs must be non-null");
    if (bath && bath->kind != TACO_J_OHMIC && bath->kind != TACO_J_DRUDE_LORENTZ)
        throw Error(TACO_ERR_INVALID, "unknown spectral density kind");
    if (bath && !(bath->beta > 0.0)) throw Error(TACO_ERR_INVALID, "beta must be > 0");
    if (steps < 1) throw Error(TACO_ERR_INVALID, "steps must be >= 1");
    if (steps + 1 > Nt) throw Error(TACO_ERR_RANGE, "steps exceed gamma_series rows");
    if (checkpoint_every < 1) throw Error(TACO_ERR_INVALID, "checkpoint_every must be >= 1");
    auto tp = std::make_unique<taco_tape>();
    tp->ctx = ctx;
    if (!ctx->tape_pool.empty()) {
        tp->bufs.swap(*ctx->tape_pool.back());
        ctx->tape_pool.pop_back();
    }
    tp->Nt = Nt;
    tp->T = steps + 1;
    tp->steps = steps;
    tp->nf = nf;
    tp->N = N;
    tp->num_ops = num_ops;
    tp->order = in.order;
    tp->K = static_cast<int>(std::min<std::size_t>(checkpoint_every, steps));
    tp->dt = in.dt;
    const std::size_t N2 = static_cast<std::size_t>(N) * N;
    tp->h_eps.assign(in.eps, in.eps + N);
    tp->h_omegas.assign(in.omegas, in.omegas + nf);
    tp->h_p2f.assign(in.pair_to_freq, in.pair_to_freq + N2);
    tp->h_mirror.assign(in.mirror, in.mirror + nf);
    const auto* oc = reinterpret_cast<const std::complex<double>*>(in.ops);
    tp->h_ops.assign(oc, oc + N2 * num_ops);
    auto put = [&](DevBuf& buf, const void* host, std::size_t bytes) {
        buf.ensure(std::max<std::size_t>(bytes, 16));
        HIP_OK(hipMemcpy(buf.p, host, bytes, hipMemcpyHostToDevice));
    };
    if (!bath) put(tp->gamma, in.gamma, Nt * nf * 16);
    put(tp->omegas, in.omegas, nf * sizeof(double));
    put(tp->mirror, in.mirror, nf * sizeof(int));
    put(tp->p2f, in.pair_to_freq, N2 * sizeof(int));
    put(tp->ops, in.ops, N2 * num_ops * 16);
    put(tp->eps, in.eps, N * sizeof(double));
    if (bath) {
        tp->bath = true;
        tp->bath_kind = bath->kind;
        tp->bath_p0 = bath->p0;
        tp->bath_p1 = bath->p1;
        tp->bath_beta = bath->beta;
        const BathLayout Ly = bath_layout(Nt, nf, tp->T);
        tp->bufs.bath.ensure(Ly.total * 16);
        tp->gamma.ensure(Nt * nf * 16);
        double2* dC = tp->bufs.bath.as<double2>();
        bcf_device(ctx, Nt - 1, in.dt, bath->kind, bath->p0, bath->p1, bath->beta, dC);
        trapz_device(ctx, dC, Nt, in.dt, in.omegas, tp->omegas.as<double>(), nf, dC + Ly.steps,
                     tp->gamma.as<double2>());
    }
    return tp;
}

// Reverse of a bath-sourced tape's bath chain: Gamma_bar = gb4 + g2 ([nf][T]
// on the device, either null for a zero part) -> dLoss/d(p0, p1, beta) into
// bath_bar[3], and the trapz Gamma's part of omega_bar added to omb.
static void tape_bath_adjoint(taco_tape& tp, const double2* gb4, const double2* g2, double* bath_bar,
                              std::vector<double>& omb) {
    taco_ctx* c = tp.ctx;
    if (!std::isfinite(tp.bath_beta))
        throw Error(TACO_ERR_INVALID, "the bath cotangents need a finite beta");
    const std::size_t T = tp.T, nfz = static_cast<std::size_t>(tp.nf);
    const BathLayout Ly = bath_layout(tp.Nt, tp.nf, T);
    double2* base = tp.bufs.bath.as<double2>();
    const double2* dC = base;
    double2* gbar = base + Ly.gbar;
    double2* Q = base + Ly.q;
    double2* x = base + Ly.x;
    double* wbar = reinterpret_cast<double*>(base + Ly.wbar);
    double* part = reinterpret_cast<double*>(base + Ly.part);
    double* out = reinterpret_cast<double*>(base + Ly.out);
    const long ng = static_cast<long>(nfz * T);
    hipLaunchKernelGGL(k_gbar_sum, dim3(static_cast<unsigned>((ng + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, gb4, g2, ng, gbar);
    hipLaunchKernelGGL(k_trapz_adj_q, dim3(static_cast<unsigned>(nfz)), dim3(kBlock), 0, c->stream, gbar,
                       static_cast<long>(T), dC, tp.dt, tp.omegas.as<double>(), Q, wbar);
    hipLaunchKernelGGL(k_trapz_adj_c, dim3(static_cast<unsigned>((Ly.nfft + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, Q, static_cast<long>(T), tp.nf, tp.dt, tp.omegas.as<double>(),
                       static_cast<long>(Ly.nfft), x);
    HIP_OK(hipGetLastError());
    const SpecOut X = fft_forward_dev(c, x, Ly.L, base + Ly.tw);
    const double domega = 2.0 * kPiHost / (tp.dt * static_cast<double>(Ly.nfft));
    BcfAdjArgs A{};
    A.spec = X.X;
    A.a = X.a;
    A.b = X.b;
    A.nfft = static_cast<long>(Ly.nfft);
    A.scale = domega / kPiHost;
    A.domega = domega;
    A.beta = tp.bath_beta;
    A.p0 = tp.bath_p0;
    A.p1 = tp.bath_p1;
    A.kind = tp.bath_kind;
    A.part = part;
    hipLaunchKernelGGL(k_bcf_adj_terms, dim3(static_cast<unsigned>(Ly.nblk)), dim3(kBlock), 0, c->stream, A);
    hipLaunchKernelGGL(k_bcf_adj_final, dim3(1), dim3(64), 0, c->stream, A, static_cast<int>(Ly.nblk), out);
    HIP_OK(hipGetLastError());
    double r[5];
    std::vector<double> wb(nfz);
    HIP_OK(hipMemcpyAsync(r, out, sizeof r, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(wb.data(), wbar, nfz * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (std::size_t b = 0; b < nfz; ++b) omb[b] += wb[b];
    // S_0 = (pi/4) J(dw) (1 + e) / (1 - e), e = exp(-beta dw); S_{nfft/2} =
    // (pi/4) J(w_Ny) (1 + 2 xN / (1 - xN)), xN = exp(-beta w_Ny)
    const double beta = tp.bath_beta, p0 = tp.bath_p0, p1 = tp.bath_p1;
    auto spec = [&](double w, double& J, double& d0, double& d1) {
        if (tp.bath_kind == TACO_J_OHMIC) {
            const double e = std::exp(-w / p1);
            J = p0 * w * e;
            d0 = w * e;
            d1 = J * w / (p1 * p1);
        } else {
            const double den = w * w + p1 * p1;
            J = 2.0 * p0 * p1 * w / den;
            d0 = 2.0 * p1 * w / den;
            d1 = 2.0 * p0 * w * (w * w - p1 * p1) / (den * den);
        }
    };
    double g0 = r[0], g1 = r[1], gb = r[2];
    double J0, a0, a1;
    spec(domega, J0, a0, a1);
    const double e = std::exp(-beta * domega);
    const double q0 = 0.5 * (kPiHost / 2.0) * (1.0 + e) / (1.0 - e);
    g0 += r[3] * q0 * a0;
    g1 += r[3] * q0 * a1;
    gb += r[3] * 0.5 * (kPiHost / 2.0) * J0 * (-2.0 * domega * e / ((1.0 - e) * (1.0 - e)));
    const double wny = kPiHost / tp.dt;
    double Jn, n0, n1;
    spec(wny, Jn, n0, n1);
    const double xN = std::exp(-beta * wny);
    const double qn = 0.25 * kPiHost * (1.0 + 2.0 * xN / (1.0 - xN));
    g0 += r[4] * qn * n0;
    g1 += r[4] * qn * n1;
    gb += r[4] * 0.25 * kPiHost * Jn * (-2.0 * wny * xN / ((1.0 - xN) * (1.0 - xN)));
    bath_bar[0] = g0;
    bath_bar[1] = g1;
    bath_bar[2] = gb;
}

// Whole slices ([3][T][nf3]: F, C, R) to and from a packed copy holding only
// the entries the L4 passes read: F at its needed triples, then C, then R,
// each [T][count]; idx holds the three ascending triple lists back to back.
// PACK copies whole -> packed, otherwise packed -> whole (the other whole
// entries are left as they are: nothing reads them).
template <bool PACK>
__global__ __launch_bounds__(kBlock) void k_slices_pack(double2* whole, double2* pk, const int* idx, long nF, long nC,
                                                          long nR, long nf3, long T) {
    const long q = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= nF + nC + nR) return;
    const int comp = q < nF ? 0 : (q < nF + nC ? 1 : 2);
    const long first = comp == 0 ? 0 : (comp == 1 ? nF : nF + nC);
    const long cnt = comp == 0 ? nF : (comp == 1 ? nC : nR);
    const long tri = idx[q];
    for (long t = blockIdx.y; t < T; t += gridDim.y) {
        const long w = (comp * T + t) * nf3 + tri;
        const long p = first * T + t * cnt + (q - first);
        if constexpr (PACK)
            pk[p] = whole[w];
        else
            whole[w] = pk[p];
    }
}

}  // namespace taco_engine

// Several tapes of one N, step count and checkpoint interval (parameter
// sets of one model): each set's F/C/R, L4 and TCL2 passes run in turn, the
// forward RK4 of every set is one launch (a workgroup per set), and so are
// the reverse RK4's recompute and walk.  The sets' F/C/R slices stay on the
// device while they fit the context's slice budget: whole when every set's
// do, else packed to the entries the L4 passes read (about a third of
// 3 nf^3 at d = 16), set by set in order while the running total fits.  The
// sets past it, and the packed sets while they are unpacked or made, share
// one buffer of one set's whole slices, and the reverse pass recomputes the
// slices of the sets past the budget (the same pass, the same values).
struct taco_mtape {
    taco_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<taco_tape>> sets;
    bool keep = true;             // every set keeps its slices
    std::vector<char> keep_set;   // per set: 0 recomputed, 1 kept whole, 2 kept packed
    std::vector<std::size_t> off; // kept sets: offset of their slices in A->slices (complex values)
    std::vector<std::size_t> pk_at;          // packed sets: offset of their triple lists in A->pkidx (ints)
    std::vector<std::array<long, 3>> pk_n;   // packed sets: F, C and R entry counts per time
    // A->samp [S][T][D] sampled states, A->rv [S][rk_rev_values] reverse RK4
    // state, A->args launch arguments, A->mid [S][steps][D^2] midpoint
    // generators of the forward RK4 (large D)
    taco_engine::MtapeArena own;
    taco_engine::MtapeArena* A = nullptr;
    ~taco_mtape() {
        if (ctx && ctx->marena_owner == this) ctx->marena_owner = nullptr;
    }
};

namespace taco_engine {

// The context's multi-set buffers when no other live multi-set tape holds
// them, else buffers of the tape's own.
static void mtape_claim(taco_mtape& mt) {
    taco_ctx* c = mt.ctx;
    if (!c->marena_owner) {
        c->marena_owner = &mt;
        mt.A = &c->marena;
    } else {
        mt.A = &mt.own;
    }
}

// Where set s's whole slices are read and written: its own region when kept
// whole, else the shared buffer.
static double2* mtape_slices(taco_mtape& mt, std::size_t s) {
    return mt.keep_set[s] == 1 ? mt.A->slices.as<double2>() + mt.off[s] : mt.A->shared.as<double2>();
}

// Packs set s's whole slices from the shared buffer into its region
// (PACK), or unpacks them back: every entry the L4 passes read is copied,
// bit for bit.
template <bool PACK>
static void mtape_pack(taco_mtape& mt, std::size_t s) {
    const taco_tape& tp = *mt.sets[s];
    const long nf3 = static_cast<long>(tp.nf) * tp.nf * tp.nf, T = static_cast<long>(tp.T);
    const std::array<long, 3>& n = mt.pk_n[s];
    const long ntot = n[0] + n[1] + n[2];
    if (ntot == 0 || T == 0) return;
    const dim3 grid(static_cast<unsigned>((ntot + kBlock - 1) / kBlock), static_cast<unsigned>(std::min<long>(T, 65535)));
    hipLaunchKernelGGL(k_slices_pack<PACK>, grid, dim3(kBlock), 0, mt.ctx->stream, mt.A->shared.as<double2>(),
                       mt.A->slices.as<double2>() + mt.off[s], mt.A->pkidx.as<const int>() + mt.pk_at[s], n[0], n[1],
                       n[2], nf3, T);
    HIP_OK(hipGetLastError());
}

// Chooses each set's slice mode within the slice budget (whole for all, else
// packed in order while they fit, the rest recomputed) and uploads the
// packed sets' triple lists, one copy per distinct bucket map.
static void mtape_plan(taco_mtape& mt) {
    const std::size_t S = mt.sets.size();
    const std::size_t budget = slice_budget(mt.ctx);
    mt.keep_set.assign(S, 1);
    mt.off.assign(S, 0);
    mt.pk_at.assign(S, 0);
    mt.pk_n.assign(S, std::array<long, 3>{0, 0, 0});
    mt.keep = true;
    std::size_t total = 0, most = 0;
    for (std::size_t s = 0; s < S; ++s) {
        const std::size_t b = tape_slice_bytes(*mt.sets[s]);
        most = std::max(most, b);
        total += b;
    }
    if (total <= budget) {
        total = 0;
        for (std::size_t s = 0; s < S; ++s) {
            mt.off[s] = total / 16;
            total += tape_slice_bytes(*mt.sets[s]);
        }
        mt.A->slices.ensure(std::max<std::size_t>(total, 16));
        return;
    }
    std::vector<int> lists;
    std::vector<std::pair<std::vector<int>, std::size_t>> seen;  // bucket-map key -> set using those lists
    total = 0;
    for (std::size_t s = 0; s < S; ++s) {
        const taco_tape& tp = *mt.sets[s];
        if (tp.order != 4) continue;  // no slices
        std::vector<int> key(tp.h_p2f);
        key.push_back(tp.N);
        key.push_back(tp.nf);
        auto it = std::find_if(seen.begin(), seen.end(), [&](const auto& e) { return e.first == key; });
        if (it != seen.end()) {
            mt.pk_at[s] = mt.pk_at[it->second];
            mt.pk_n[s] = mt.pk_n[it->second];
        } else {
            const std::vector<unsigned char> need = needed_triples(tp.h_p2f.data(), tp.N, tp.nf);
            mt.pk_at[s] = lists.size();
            for (int bit = 0; bit < 3; ++bit) {
                long cnt = 0;
                for (std::size_t t = 0; t < need.size(); ++t)
                    if (need[t] & (1 << bit)) {
                        lists.push_back(static_cast<int>(t));
                        ++cnt;
                    }
                mt.pk_n[s][bit] = cnt;
            }
            seen.emplace_back(std::move(key), s);
        }
        const std::size_t b =
            static_cast<std::size_t>(mt.pk_n[s][0] + mt.pk_n[s][1] + mt.pk_n[s][2]) * tp.T * 16;
        if (b <= budget - std::min(budget, total)) {
            mt.keep_set[s] = 2;
            mt.off[s] = total / 16;
            total += b;
        } else {
            mt.keep_set[s] = 0;
            mt.keep = false;
        }
    }
    mt.A->slices.ensure(std::max<std::size_t>(total, 16));
    mt.A->shared.ensure(std::max<std::size_t>(most, 16));
    mt.A->pkidx.ensure(std::max<std::size_t>(lists.size(), 1) * sizeof(int));
    if (!lists.empty()) {
        // in stream order after the arena's earlier users; lists dies on return
        HIP_OK(hipMemcpyAsync(mt.A->pkidx.p, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice,
                              mt.ctx->stream));
        HIP_OK(hipStreamSynchronize(mt.ctx->stream));
    }
}

static void mtape_forward(taco_mtape& mt, const taco_tape_set* in, double* const* samples_out) {
    taco_ctx* c = mt.ctx;
    const std::size_t S = mt.sets.size();
    const taco_tape& t0 = *mt.sets[0];
    const std::size_t D = static_cast<std::size_t>(t0.N) * t0.N, T = t0.T;
    mtape_plan(mt);
    for (std::size_t s = 0; s < S; ++s) {
        tape_generator(*mt.sets[s], mtape_slices(mt, s));
        if (mt.keep_set[s] == 2) mtape_pack<true>(mt, s);
    }
    // one RK4 launch over every set: r0 in, samples out ([S][T][D]), final r after them
    mt.A->samp.ensure((S * T + 2 * S) * D * 16);
    double2* samp = mt.A->samp.as<double2>();
    double2* r0d = samp + S * T * D;
    double2* rend = r0d + S * D;
    for (std::size_t s = 0; s < S; ++s)
        HIP_OK(hipMemcpyAsync(r0d + s * D, in[s].r0, D * 16, hipMemcpyHostToDevice, c->stream));
    const int Di = static_cast<int>(D);
    const int rows_thr = rk4_rows_threads(Di);
    const int nstage = rows_thr > 0 ? 0 : rk4_nstage(Di, 6, static_cast<long>(t0.steps) + 1);
    const std::size_t lds = rows_thr > 0 ? static_cast<std::size_t>(3) * D * 16 : rk4_lds(Di, 6, nstage);
    if (lds > kRkLds) throw Error(TACO_ERR_INVALID, "RK4 kernel supports D <= 1706");
    std::vector<Rk4Sys> sys(S);
    const bool mids = rows_thr == 0 && nstage < 2 && rk4_midpoints_fit(Di, t0.steps * S);
    if (mids) mt.A->mid.ensure(S * t0.steps * D * D * 16);
    for (std::size_t s = 0; s < S; ++s) {
        const taco_tape& tp = *mt.sets[s];
        sys[s] = Rk4Sys{tp.lgen.as<double2>(), nstage, r0d + s * D, rend + s * D, samp + s * T * D,
                        static_cast<long>(tp.steps), tp.dt, Di};
        if (mids) {
            const long D2 = static_cast<long>(D) * D, total = static_cast<long>(tp.steps) * D2;
            double2* m = mt.A->mid.as<double2>() + s * tp.steps * D * D;
            hipLaunchKernelGGL(k_midpoints, dim3(static_cast<unsigned>(std::min<long>((total + kBlock - 1) / kBlock,
                                                                                     65536))),
                               dim3(kBlock), 0, c->stream, tp.lgen.as<double2>(), static_cast<long>(tp.steps), D2, m);
            sys[s].Lh = m;
        }
    }
    mt.A->args.ensure(std::max(S * sizeof(Rk4Sys), 2 * S * sizeof(Rk4AdjArgs)));
    HIP_OK(hipMemcpyAsync(mt.A->args.p, sys.data(), S * sizeof(Rk4Sys), hipMemcpyHostToDevice, c->stream));
    const double rk_bytes = static_cast<double>(S) * ((t0.steps + 1) * D * D * 16.0 + (t0.steps + 1) * D * 16.0);
    timed(c, K_RK4, rk_bytes, [&] {
        hipLaunchKernelGGL(k_rk4_batch, dim3(static_cast<unsigned>(S)), dim3(rows_thr > 0 ? rows_thr : rk4_block(Di)),
                           lds, c->stream, mt.A->args.as<const Rk4Sys>());
    });
    HIP_OK(hipGetLastError());
    for (std::size_t s = 0; s < S; ++s) tape_checkpoints(*mt.sets[s], samp + s * T * D);
    for (std::size_t s = 0; s < S; ++s)
        if (samples_out && samples_out[s])
            HIP_OK(hipMemcpyAsync(samples_out[s], samp + s * T * D, T * D * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
}

static void mtape_adjoint(taco_mtape& mt, const double* const* samples_bar, double* const* gamma_bar,
                          double* const* omega_bar, double* const* ops_bar, double* const* eps_bar,
                          double* const* r0_bar, double* const* bath_bar = nullptr) {
    taco_ctx* c = mt.ctx;
    const std::size_t S = mt.sets.size();
    const taco_tape& t0 = *mt.sets[0];
    const std::size_t D = static_cast<std::size_t>(t0.N) * t0.N, D2 = D * D, T = t0.T;
    const std::size_t per = rk_rev_values(t0);
    mt.A->rv.ensure(S * per * 16);
    std::vector<RkRev> rv(S);
    std::vector<Rk4AdjArgs> args(2 * S);
    RkRevShape sh{};
    for (std::size_t s = 0; s < S; ++s) {
        rv[s] = rk_rev_at(*mt.sets[s], mt.A->rv.as<double2>() + s * per);
        HIP_OK(hipMemcpyAsync(rv[s].gb, samples_bar[s], T * D * 16, hipMemcpyHostToDevice, c->stream));
        sh = rk_rev_shape(*mt.sets[s], rv[s]);
        args[s] = sh.Rf;
        args[S + s] = sh.R;
        if (rv[s].LsH) {
            const int tiles = (static_cast<int>(D) + 15) / 16;
            hipLaunchKernelGGL(k_conj_transpose, dim3(static_cast<unsigned>(tiles * tiles), static_cast<unsigned>(T)),
                               dim3(kBlock), 0, c->stream, mt.sets[s]->lgen.as<double2>(), rv[s].LsH,
                               static_cast<int>(D), static_cast<long>(T));
            const long D2l = static_cast<long>(D2), total = static_cast<long>(T - 1) * D2l;
            hipLaunchKernelGGL(k_midpoints,
                               dim3(static_cast<unsigned>(std::min<long>((total + kBlock - 1) / kBlock, 65536))),
                               dim3(kBlock), 0, c->stream, static_cast<const double2*>(rv[s].LsH),
                               static_cast<long>(T - 1), D2l, rv[s].LhH);
        } else if (!sh.carry) {
            const long cntL = static_cast<long>(T * D2);
            hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((cntL + kBlock - 1) / kBlock)), dim3(kBlock),
                               0, c->stream, rv[s].Lb, cntL);
        }
    }
    mt.A->args.ensure(std::max(S * sizeof(Rk4Sys), 2 * S * sizeof(Rk4AdjArgs)));
    const Rk4AdjArgs* dargs = mt.A->args.as<const Rk4AdjArgs>();
    HIP_OK(hipMemcpyAsync(mt.A->args.p, args.data(), 2 * S * sizeof(Rk4AdjArgs), hipMemcpyHostToDevice, c->stream));
    const int Di = static_cast<int>(D);
    timed(c, K_RK4, sh.bytes * static_cast<double>(S), [&] {
        if (sh.rows_f) {
            rows_dispatch(Di, [&](auto dc) {
                constexpr int DT = decltype(dc)::value;
                hipLaunchKernelGGL(k_rk4_recompute_rows_multi<DT>,
                                   dim3(static_cast<unsigned>(sh.nseg), static_cast<unsigned>(S)),
                                   dim3(rk4_rows_threads(DT)), 3 * DT * 16, c->stream, dargs);
            });
        } else {
            hipLaunchKernelGGL(k_rk4_recompute_multi, dim3(static_cast<unsigned>(sh.nseg), static_cast<unsigned>(S)),
                               dim3(kBlock), sh.lds_f, c->stream, dargs);
        }
        if (sh.rows_b) {
            rows_dispatch(Di, [&](auto dc) {
                constexpr int DT = decltype(dc)::value;
                if constexpr (DT <= 16)
                    hipLaunchKernelGGL(k_rk4_adjoint_rows_multi<DT>, dim3(static_cast<unsigned>(S)), dim3(64),
                                       5 * DT * 16, c->stream, dargs + S);
            });
        } else {
            hipLaunchKernelGGL(k_rk4_adjoint_multi, dim3(static_cast<unsigned>(S)), dim3(kBlock), sh.lds_b, c->stream,
                               dargs + S);
        }
        if (rv[0].vec)
            hipLaunchKernelGGL(k_rk4_lbar_multi, dim3(static_cast<unsigned>((D2 + kBlock - 1) / kBlock),
                                                      static_cast<unsigned>(T), static_cast<unsigned>(S)),
                               dim3(kBlock), 0, c->stream, dargs + S);
    });
    HIP_OK(hipGetLastError());
    for (std::size_t s = 0; s < S; ++s) {
        taco_tape& tp = *mt.sets[s];
        if (tp.order == 4 && mt.keep_set[s] == 0) tape_slices(tp, mt.A->shared.as<double2>());
        if (tp.order == 4 && mt.keep_set[s] == 2) mtape_pack<false>(mt, s);
        auto pick = [&](double* const* v) { return v ? v[s] : nullptr; };
        tape_adjoint_tail(tp, rv[s], mtape_slices(mt, s), pick(gamma_bar), pick(omega_bar), pick(ops_bar),
                          pick(eps_bar), pick(r0_bar), pick(bath_bar));
    }
}

}  // namespace taco_engine

namespace taco_engine {
// ---------------------------------------------------------------------------
// System preparation on the device (system.hpp:37-157, tcl4.cpp:40-101): a
// workgroup per Hamiltonian runs the Jacobi sweeps of system_prep.hpp with
// each rotation's independent row and column updates spread over its
// threads, then the ordering, the gauge, the buckets with mirror and zero
// bucket, and A_eig = U^H A U; the host builder runs the same steps, so the
// results are equal bit for bit.
// ---------------------------------------------------------------------------
constexpr int kSysMaxN = 64;

struct SysArgs {
    const double2* H;     // [nsys][N*N] column-major
    const double2* Alab;  // [nsys][num_ops][N*N]
    int N, num_ops;
    double tol;
    double* eps;          // [nsys][N]
    double2* U;           // [nsys][N*N]
    double2* Aeig;        // [nsys][num_ops][N*N]
    int* nf;              // [nsys]
    double* omegas;       // [nsys][N*N]
    int* mirror;          // [nsys][N*N]
    int* p2f;             // [nsys][N*N]
    int* zero;            // [nsys]
    double* rep;          // scratch [nsys][N*N]
    int* idx;             // scratch [nsys][3][N*N]
};

__global__ __launch_bounds__(kBlock) void k_system_build(SysArgs S) {
    using namespace taco_sys;
    extern __shared__ double2 shs[];
    __shared__ int flag;
    __shared__ int order[kSysMaxN];
    const int N = S.N;
    const long NN = static_cast<long>(N) * N;
    const long sys = blockIdx.x;
    Cx* A = reinterpret_cast<Cx*>(shs);
    Cx* V = A + NN;
    const double2* H = S.H + sys * NN;
    for (long e = threadIdx.x; e < NN; e += blockDim.x) {
        A[e] = cx(H[e].x, H[e].y);
        V[e] = cx((e % N) == (e / N) ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    for (int sweep = 0; sweep < 100; ++sweep) {
        if (threadIdx.x == 0) flag = converged(A, N) ? 1 : 0;
        __syncthreads();
        if (flag) break;
        for (int p = 0; p < N; ++p)
            for (int q = p + 1; q < N; ++q) {
                const Rot g = rotation(A, N, p, q);  // the same inputs in every thread
                __syncthreads();
                if (g.skip) continue;
                for (int r = threadIdx.x; r < N; r += blockDim.x) rot_cols(A, N, p, q, g, r);
                __syncthreads();
                for (int c = threadIdx.x; c < N; c += blockDim.x) rot_rows(A, N, p, q, g, c);
                __syncthreads();
                if (threadIdx.x == 0) rot_fix(A, N, p, q);
                for (int r = threadIdx.x; r < N; r += blockDim.x) rot_cols(V, N, p, q, g, r);
                __syncthreads();
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) sort_diag(A, N, order);
    __syncthreads();
    double* eps = S.eps + sys * N;
    Cx* U = reinterpret_cast<Cx*>(S.U + sys * NN);
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        eps[c] = A[static_cast<long>(order[c]) * (N + 1)].re;
        gauge_column(V, N, order[c], U + static_cast<long>(c) * N);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int* idx = S.idx + sys * 3 * NN;
        S.nf[sys] = buckets(eps, N, S.tol, S.rep + sys * NN, idx, idx + NN, idx + 2 * NN, S.omegas + sys * NN,
                            S.mirror + sys * NN, S.p2f + sys * NN, S.zero + sys);
    }
    // A_eig = U^H A U, the product U^H A staged in the LDS rows A used
    for (int o = 0; o < S.num_ops; ++o) {
        const Cx* Al = reinterpret_cast<const Cx*>(S.Alab + (sys * S.num_ops + o) * NN);
        for (long e = threadIdx.x; e < NN; e += blockDim.x)
            A[e] = uha(U, Al, N, static_cast<int>(e % N), static_cast<int>(e / N));
        __syncthreads();
        double2* E = S.Aeig + (sys * S.num_ops + o) * NN;
        for (long e = threadIdx.x; e < NN; e += blockDim.x) {
            const Cx v = tu(A, U, N, static_cast<int>(e % N), static_cast<int>(e / N));
            E[e] = make_double2(v.re, v.im);
        }
        __syncthreads();
    }
}
}  // namespace taco_engine

using namespace taco_engine;

extern "C" {

const char* taco_last_error(void) { return g_last_error.c_str(); }
const char* taco_version(void) { return "taco-mi355x 0.1.0 (gfx950)"; }

void taco_set_fcr_fft_pad_factor(size_t factor) { g_pad_factor = factor; }
size_t taco_get_fcr_fft_pad_factor(void) { return g_pad_factor; }

int taco_set_phase_mode(int mode) {
    return guarded([&] { set_phase_mode(mode); });
}
int taco_get_phase_mode(void) { return phase_mode(); }

int taco_set_fft_mode(int mode) {
    return guarded([&] { set_fft_mode(mode); });
}
int taco_get_fft_mode(void) { return fft_mode(); }

int taco_set_slice_budget(taco_ctx* ctx, size_t bytes) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->slice_budget = bytes;
    });
}
size_t taco_get_slice_budget(const taco_ctx* ctx) { return ctx ? slice_budget(ctx) : 0; }

int taco_ctx_create(int device, taco_ctx** out) {
    return guarded([&] {
        if (!out) throw Error(TACO_ERR_INVALID, "null output pointer");
        int count = 0;
        HIP_OK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Error(TACO_ERR_INVALID, "gpu_id out of range");
        DeviceGuard g(device);
        auto c = std::make_unique<taco_ctx>();
        c->device = device;
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_crtab), cr_table(), sizeof(double) * 4 * taco_cr::kTabN, 0,
                                 hipMemcpyHostToDevice));
        HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_exptab), exp_table(), sizeof(double) * taco_cr::kExpTabSize, 0,
                                 hipMemcpyHostToDevice));
        HIP_OK(hipEventCreate(&c->ev0));
        HIP_OK(hipEventCreate(&c->ev1));
        for (const void* fn : {reinterpret_cast<const void*>(k_rk4), reinterpret_cast<const void*>(k_rk4_batch),
                               reinterpret_cast<const void*>(k_rk4_adjoint),
                               reinterpret_cast<const void*>(k_rk4_recompute),
                               reinterpret_cast<const void*>(k_rk4_adjoint_multi),
                               reinterpret_cast<const void*>(k_rk4_recompute_multi)})
            HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kRkLds)));
        *out = c.release();
    });
}

int taco_ctx_destroy(taco_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        {
            DeviceGuard g(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipEventDestroy(ctx->ev0);
            (void)hipEventDestroy(ctx->ev1);
            (void)hipStreamDestroy(ctx->stream);
            ctx->tw.clear();
        }
        DeviceGuard g2(ctx->device);
        delete ctx;
    });
}

void* taco_ctx_stream(taco_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int taco_profile_enable(taco_ctx* ctx, int on) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->prof.resolve();
        ctx->prof.on = on != 0;
    });
}

int taco_profile_reset(taco_ctx* ctx) {
    return guarded([&] {
        check_ctx(ctx);
        DeviceGuard g(ctx->device);
        ctx->prof.reset();
    });
}

int taco_profile_get(taco_ctx* ctx, int kernel, double* total_ms, long* launches, double* algo_bytes) {
    return guarded([&] {
        check_ctx(ctx);
        if (kernel < 0 || kernel >= K_COUNT) throw Error(TACO_ERR_RANGE, "unknown kernel id");
        DeviceGuard g(ctx->device);
        ctx->prof.resolve();
        if (total_ms) *total_ms = ctx->prof.ms[kernel];
        if (launches) *launches = ctx->prof.launches[kernel];
        if (algo_bytes) *algo_bytes = ctx->prof.bytes[kernel];
    });
}

int taco_fft_tile_choice(taco_ctx* ctx, int log2_nfft, int* cfg3) {
    return guarded([&] {
        check_ctx(ctx);
        auto it = ctx->fft_choice.find(log2_nfft);
        for (int k = 0; k < 3; ++k) cfg3[k] = it == ctx->fft_choice.end() ? -1 : it->second[k];
    });
}

int taco_cr_sincos(taco_ctx* ctx, const double* x, size_t n, double* s, double* c) {
    return guarded([&] {
        if (!ctx) {
            cr_sincos_host(x, n, s, c);
            return;
        }
        if (n == 0) return;
        DeviceGuard g(ctx->device);
        ctx->misc.ensure(3 * n * sizeof(double));
        double* dx = ctx->misc.as<double>();
        HIP_OK(hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_cr_sincos, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream, dx,
                           static_cast<long>(n), dx + n, dx + 2 * n);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(s, dx + n, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(c, dx + 2 * n, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}


int taco_cr_exp(taco_ctx* ctx, const double* x, size_t n, double* out) {
    return guarded([&] {
        if (!ctx) {
            cr_exp_host(x, n, out);
            return;
        }
        if (n == 0) return;
        DeviceGuard g(ctx->device);
        ctx->misc.ensure(2 * n * sizeof(double));
        double* dx = ctx->misc.as<double>();
        HIP_OK(hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_cr_exp, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream, dx,
                           static_cast<long>(n), dx + n);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out, dx + n, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_bath_gamma_hip(taco_ctx* ctx, size_t N, double dt, int kind, double p0, double p1, double beta,
                        const double* omegas, int nf, double* C_out, double* gamma_out) {
    return guarded([&] {
        check_ctx(ctx);
        if (nf < 0) throw Error(TACO_ERR_INVALID, "nf must be >= 0");
        if (nf > 0 && (!omegas || !gamma_out)) throw Error(TACO_ERR_INVALID, "omegas and gamma_out are required");
        DeviceGuard g(ctx->device);
        const std::size_t Nt = N + 1;
        ctx->inter2.ensure((Nt + static_cast<std::size_t>(nf) * Nt + nf) * 16 + static_cast<std::size_t>(nf + 1) * 8);
        double2* dC = ctx->inter2.as<double2>();
        double2* dG = dC + Nt;
        double2* d_steps = dG + static_cast<std::size_t>(nf) * Nt;
        double* d_om = reinterpret_cast<double*>(d_steps + nf);
        bcf_device(ctx, N, dt, kind, p0, p1, beta, dC);
        if (nf > 0) {
            HIP_OK(hipMemcpyAsync(d_om, omegas, static_cast<std::size_t>(nf) * 8, hipMemcpyHostToDevice, ctx->stream));
            trapz_device(ctx, dC, Nt, dt, omegas, d_om, nf, d_steps, dG);
            HIP_OK(hipMemcpyAsync(gamma_out, dG, static_cast<std::size_t>(nf) * Nt * 16, hipMemcpyDeviceToHost,
                                  ctx->stream));
        }
        if (C_out) HIP_OK(hipMemcpyAsync(C_out, dC, Nt * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_gamma_trapz_hip(taco_ctx* ctx, const double* C, size_t Nc, double dt, const double* omegas, int nf,
                         double* gamma_out) {
    return guarded([&] {
        check_ctx(ctx);
        if (nf < 0) throw Error(TACO_ERR_INVALID, "nf must be >= 0");
        if (Nc == 0 || nf == 0) return;
        DeviceGuard g(ctx->device);
        ctx->inter2.ensure((Nc + static_cast<std::size_t>(nf) * Nc + nf) * 16 + static_cast<std::size_t>(nf + 1) * 8);
        double2* dC = ctx->inter2.as<double2>();
        double2* dG = dC + Nc;
        double2* d_steps = dG + static_cast<std::size_t>(nf) * Nc;
        double* d_om = reinterpret_cast<double*>(d_steps + nf);
        HIP_OK(hipMemcpyAsync(dC, C, Nc * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d_om, omegas, static_cast<std::size_t>(nf) * 8, hipMemcpyHostToDevice, ctx->stream));
        trapz_device(ctx, dC, Nc, dt, omegas, d_om, nf, d_steps, dG);
        HIP_OK(hipMemcpyAsync(gamma_out, dG, static_cast<std::size_t>(nf) * Nc * 16, hipMemcpyDeviceToHost,
                              ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_system_build(const double* H, int N, const double* A_lab, int num_ops, double freq_tol, double* eps,
                      double* U, double* A_eig, int* nf, double* omegas, int* mirror, int* pair_to_freq,
                      int* zero_index) {
    return guarded([&] {
        system_build(H, N, A_lab, num_ops, freq_tol, eps, U, A_eig, nf, omegas, mirror, pair_to_freq,
                     zero_index);
    });
}

int taco_system_build_hip(taco_ctx* ctx, int nsys, const double* H, int N, const double* A_lab, int num_ops,
                          double freq_tol, double* eps, double* U, double* A_eig, int* nf, double* omegas,
                          int* mirror, int* pair_to_freq, int* zero_index) {
    return guarded([&] {
        check_ctx(ctx);
        if (nsys < 0) throw Error(TACO_ERR_INVALID, "nsys must be >= 0");
        if (N <= 0) throw Error(TACO_ERR_INVALID, "Eigensystem: Hamiltonian dimension must be > 0");
        if (N > kSysMaxN) throw Error(TACO_ERR_INVALID, "device system build supports N <= 64");
        if (num_ops < 0) throw Error(TACO_ERR_INVALID, "system_build: num_ops must be >= 0");
        if (nsys == 0) return;
        if (!H || !eps || !U || !nf || !omegas || !mirror || !pair_to_freq || !zero_index ||
            (num_ops > 0 && (!A_lab || !A_eig)))
            throw Error(TACO_ERR_INVALID, "null array");
        const std::size_t NN = static_cast<std::size_t>(N) * N, S = static_cast<std::size_t>(nsys);
        const auto* Hc = reinterpret_cast<const std::complex<double>*>(H);
        for (std::size_t s = 0; s < S; ++s)
            for (std::size_t r = 0; r < static_cast<std::size_t>(N); ++r)
                for (std::size_t c = 0; c < static_cast<std::size_t>(N); ++c) {
                    const auto a = Hc[s * NN + r + c * N], b = Hc[s * NN + c + r * N];
                    if (std::abs(a - std::conj(b)) > 1e-12 * (1.0 + std::abs(a)) ||
                        (r == c && std::abs(a.imag()) > 1e-12 * (1.0 + std::abs(a.real()))))
                        throw Error(TACO_ERR_INVALID, "Eigensystem: Hamiltonian must be Hermitian");
                }
        DeviceGuard g(ctx->device);
        const std::size_t nops = static_cast<std::size_t>(num_ops);
        // device layout in misc: H, A_lab, U, A_eig (complex); eps, omegas, rep (double); ints
        const std::size_t nc = S * NN * (2 + 2 * nops), nd = S * (N + 2 * NN), ni = S * (2 + 5 * NN);
        ctx->misc.ensure(nc * 16 + nd * 8 + ni * 4 + 64);
        double2* dH = ctx->misc.as<double2>();
        double2* dA = dH + S * NN;
        double2* dU = dA + S * nops * NN;
        double2* dE = dU + S * NN;
        double* deps = reinterpret_cast<double*>(dE + S * nops * NN);
        double* dom = deps + S * N;
        double* drep = dom + S * NN;
        int* dnf = reinterpret_cast<int*>(drep + S * NN);
        int* dzero = dnf + S;
        int* dmir = dzero + S;
        int* dp2f = dmir + S * NN;
        int* didx = dp2f + S * NN;
        HIP_OK(hipMemcpyAsync(dH, H, S * NN * 16, hipMemcpyHostToDevice, ctx->stream));
        if (nops) HIP_OK(hipMemcpyAsync(dA, A_lab, S * nops * NN * 16, hipMemcpyHostToDevice, ctx->stream));
        SysArgs A{dH, dA, N, num_ops, std::max(freq_tol, std::numeric_limits<double>::epsilon()), deps, dU, dE, dnf,
                  dom, dmir, dp2f, dzero, drep, didx};
        const std::size_t lds = 2 * NN * 16;
        set_fn_lds(reinterpret_cast<const void*>(k_system_build), lds);
        hipLaunchKernelGGL(k_system_build, dim3(static_cast<unsigned>(S)), dim3(kBlock), lds, ctx->stream, A);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(eps, deps, S * N * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(U, dU, S * NN * 16, hipMemcpyDeviceToHost, ctx->stream));
        if (nops) HIP_OK(hipMemcpyAsync(A_eig, dE, S * nops * NN * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(nf, dnf, S * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(omegas, dom, S * NN * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(mirror, dmir, S * NN * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(pair_to_freq, dp2f, S * NN * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(zero_index, dzero, S * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_bcf_fft(size_t N, double dt, int kind, double p0, double p1, double beta, double* C_out) {
    return guarded([&] { bcf_fft(N, dt, kind, p0, p1, beta, C_out); });
}

int taco_gamma_trapz(const double* C, size_t Nc, double dt, const double* omegas, int nf, double* gamma_out) {
    return guarded([&] { gamma_trapz(C, Nc, dt, omegas, nf, gamma_out); });
}

int taco_triple_kernels(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                        const int* mirror, double dt, int method, size_t keep, double* F, double* C, double* R) {
    return guarded([&] {
        check_ctx(ctx);
        if (method != TACO_FCR_CONVOLUTION && method != TACO_FCR_DIRECT)
            throw Error(TACO_ERR_INVALID, "compute_triple_kernels: unknown FCRMethod");
        if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "dt must be > 0");
        if (nf <= 0 || Nt == 0) throw Error(TACO_ERR_INVALID, "empty gamma_series");
        for (int b = 0; b < nf; ++b)
            if (mirror[b] < 0 || mirror[b] >= nf) throw Error(TACO_ERR_INVALID, "mirror index out of range");
        if (keep == 0) keep = Nt;
        if (keep > Nt) throw Error(TACO_ERR_RANGE, "keep exceeds gamma_series rows");
        DeviceGuard g(ctx->device);
        const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
        const double2* dg = upload(ctx, ctx->gamma, reinterpret_cast<const double2*>(gamma), Nt * nf);
        const double* dom = upload(ctx, ctx->omegas, omegas, static_cast<std::size_t>(nf));
        const int* dmi = upload(ctx, ctx->mirror, mirror, static_cast<std::size_t>(nf));
        ctx->slices.ensure(3 * nf3 * keep * 16);
        CombineArgs out{};
        out.out_mode = 1;
        out.t2pos = nullptr;
        out.nf3 = static_cast<long>(nf3);
        double2* sl = ctx->slices.as<double2>();
        out.Fo = sl;
        out.Co = sl + nf3 * keep;
        out.Ro = sl + 2 * nf3 * keep;
        FcrPlan pl{dg, Nt, nf, dom, omegas, dmi, dt, keep, nullptr, 0, method};
        run_fcr(ctx, pl, out);
        HIP_OK(hipMemcpyAsync(F, out.Fo, nf3 * keep * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(C, out.Co, nf3 * keep * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(R, out.Ro, nf3 * keep * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_l4_batch(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                  const int* pair_to_freq, int N, const double* ops, int num_ops, double dt, int method,
                  const size_t* tidx, size_t ntau, double* L4_out, double* fcr_ms) {
    return guarded([&] {
        check_ctx(ctx);
        if (fcr_ms) *fcr_ms = 0.0;
        validate_system(Nt, nf, mirror, pair_to_freq, N, num_ops, dt, method);
        DeviceGuard g(ctx->device);
        const std::size_t N2 = static_cast<std::size_t>(N) * N, N4 = N2 * N2;
        const double2* dg = upload(ctx, ctx->gamma, reinterpret_cast<const double2*>(gamma), Nt * nf);
        const double* dom = upload(ctx, ctx->omegas, omegas, static_cast<std::size_t>(nf));
        const int* dmi = upload(ctx, ctx->mirror, mirror, static_cast<std::size_t>(nf));
        const int* dpf = upload(ctx, ctx->p2f, pair_to_freq, N2);
        const double2* dops = upload(ctx, ctx->ops, reinterpret_cast<const double2*>(ops), N2 * num_ops);
        TimePlan tp = l4_core(ctx, dg, Nt, nf, dom, omegas, dmi, dpf, pair_to_freq, N, dops, num_ops, dt, tidx, ntau,
                              fcr_ms, method);
        if (tp.uniq.empty()) return;
        std::vector<double> host(N4 * tp.uniq.size() * 2);
        HIP_OK(hipMemcpyAsync(host.data(), ctx->l4.p, host.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        for (std::size_t q = 0; q < tp.req2u.size(); ++q)
            std::memcpy(L4_out + q * N4 * 2, host.data() + tp.req2u[q] * N4 * 2, N4 * 16);
    });
}

int taco_l4_batch_dev(taco_ctx* ctx, const void* d_gamma, size_t Nt, int nf, const void* d_omegas,
                      const void* d_mirror, const void* d_pair_to_freq, int N, const void* d_ops, int num_ops,
                      double dt, int method, const size_t* tidx, size_t ntau, void* d_L4_out, double* fcr_ms) {
    return guarded([&] {
        check_ctx(ctx);
        if (fcr_ms) *fcr_ms = 0.0;
        validate_system(Nt, nf, nullptr, nullptr, N, num_ops, dt, method);
        DeviceGuard g(ctx->device);
        const std::size_t N2 = static_cast<std::size_t>(N) * N, N4 = N2 * N2;
        // the kernels index with these tables: validate them on the host
        std::vector<int> hm(static_cast<std::size_t>(nf)), hp(N2);
        std::vector<double> hom(static_cast<std::size_t>(nf));
        HIP_OK(hipMemcpy(hm.data(), d_mirror, hm.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(hp.data(), d_pair_to_freq, hp.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(hom.data(), d_omegas, hom.size() * sizeof(double), hipMemcpyDeviceToHost));
        validate_system(Nt, nf, hm.data(), hp.data(), N, num_ops, dt, method);
        TimePlan tp = l4_core(ctx, static_cast<const double2*>(d_gamma), Nt, nf, static_cast<const double*>(d_omegas),
                              hom.data(), static_cast<const int*>(d_mirror), static_cast<const int*>(d_pair_to_freq),
                              hp.data(), N,
                              static_cast<const double2*>(d_ops), num_ops, dt, tidx, ntau, fcr_ms, method);
        if (tp.uniq.empty()) return;
        double2* out = static_cast<double2*>(d_L4_out);
        const bool identity = tp.req2u.size() == tp.uniq.size() &&
                              std::is_sorted(tp.req2u.begin(), tp.req2u.end());
        if (identity) {
            HIP_OK(hipMemcpyAsync(out, ctx->l4.p, N4 * tp.uniq.size() * 16, hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            for (std::size_t q = 0; q < tp.req2u.size(); ++q)
                HIP_OK(hipMemcpyAsync(out + q * N4, ctx->l4.as<double2>() + tp.req2u[q] * N4, N4 * 16,
                                      hipMemcpyDeviceToDevice, ctx->stream));
        }
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_l4_single(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                   const int* pair_to_freq, int N, const double* ops, int num_ops, double dt, int method,
                   size_t time_index, double* L4_out) {
    if (time_index >= Nt) {
        set_last_error("build_TCL4_generator_hip_fused: time_index out of range");
        return TACO_ERR_RANGE;
    }
    const size_t t = time_index;
    return taco_l4_batch(ctx, gamma, Nt, nf, omegas, mirror, pair_to_freq, N, ops, num_ops, dt, method, &t, 1,
                         L4_out, nullptr);
}

int taco_l4_from_slices(taco_ctx* ctx, const double* Fs, const double* Cs, const double* Rs, const int* pair_to_freq,
                        int N, int nf, const double* ops, int num_ops, double* L4_out, double* GW_out) {
    return guarded([&] {
        check_ctx(ctx);
        if (N <= 0 || nf <= 0) throw Error(TACO_ERR_INVALID, "N and nf must be > 0");
        if (num_ops <= 0) throw Error(TACO_ERR_INVALID, "assemble_liouvillian: coupling_ops must be non-empty");
        for (int e = 0; e < N * N; ++e) {
            if (pair_to_freq[e] < 0)
                throw Error(TACO_ERR_RUNTIME, "pair_to_freq contains -1 for a required pair");
            if (pair_to_freq[e] >= nf) throw Error(TACO_ERR_INVALID, "pair_to_freq entry >= nf");
        }
        DeviceGuard g(ctx->device);
        const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
        const std::size_t N2 = static_cast<std::size_t>(N) * N, N4 = N2 * N2;
        ctx->slices.ensure(3 * nf3 * 16);
        double2* sl = ctx->slices.as<double2>();
        HIP_OK(hipMemcpyAsync(sl, Fs, nf3 * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(sl + nf3, Cs, nf3 * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(sl + 2 * nf3, Rs, nf3 * 16, hipMemcpyHostToDevice, ctx->stream));
        const int* dpf = upload(ctx, ctx->p2f, pair_to_freq, N2);
        const double2* dops = upload(ctx, ctx->ops, reinterpret_cast<const double2*>(ops), N2 * num_ops);
        ctx->l4.ensure(N4 * 16);
        ctx->gw.ensure(N4 * 16);
        L4Args A{};
        A.Fs = sl;
        A.Cs = sl + nf3;
        A.Rs = sl + 2 * nf3;
        A.p2f = dpf;
        A.ops = dops;
        A.N = N;
        A.nf = nf;
        A.num_ops = num_ops;
        A.nf3 = static_cast<long>(nf3);
        A.L4 = ctx->l4.as<double2>();
        A.GW = ctx->gw.as<double2>();
        launch_l4(ctx, A, 1);
        HIP_OK(hipMemcpyAsync(L4_out, A.L4, N4 * 16, hipMemcpyDeviceToHost, ctx->stream));
        if (GW_out) HIP_OK(hipMemcpyAsync(GW_out, A.GW, N4 * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"

// Scratch of the allocation-free entry point: the factorised L4's per-time
// intermediates, in a file-scope device array that each GPU holds from
// module load to process end.  Host code must not pass `scratch` itself to a
// kernel (on the host it only names the symbol): it uses the device address
// that hipGetSymbolAddress returns, resolved once per device.
namespace taco_engine {
constexpr std::size_t kNoallocScratchBytes = std::size_t{64} << 20;
static __device__ __align__(256) double2 scratch[kNoallocScratchBytes / sizeof(double2)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double2* g_scratch_ptr[kMaxDevices] = {};

static double2* scratch_on_current_device() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(TACO_ERR_RUNTIME, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(TACO_ERR_RUNTIME, "scratch symbol has no device address");
        g_scratch_ptr[dev] = static_cast<double2*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace taco_engine

extern "C" {

int taco_l4_from_slices_noalloc(const void* Fs, const void* Cs, const void* Rs, const void* pair_to_freq, int N,
                                int nf, const void* ops, int num_ops, size_t npos, void* L4_out, void* GW_out,
                                void* stream) {
    return guarded([&] {
        if (N <= 0 || nf <= 0) throw Error(TACO_ERR_INVALID, "N and nf must be > 0");
        if (num_ops <= 0) throw Error(TACO_ERR_INVALID, "assemble_liouvillian: coupling_ops must be non-empty");
        if (npos == 0) return;
        if (!Fs || !Cs || !Rs || !pair_to_freq || !ops || !L4_out)
            throw Error(TACO_ERR_INVALID, "device pointers must be non-null");
        const long N2 = static_cast<long>(N) * N;
        const std::size_t lds = static_cast<std::size_t>(num_ops) * N2 * 16 + static_cast<std::size_t>(N2) * 4;
        if (lds > 64 * 1024) throw Error(TACO_ERR_INVALID, "factorised L4 kernel: coupling operators exceed 64 KiB of LDS");
        const long per = l4f_per_pos(N, nf, num_ops);
        const std::size_t per_bytes = static_cast<std::size_t>(per) * sizeof(double2);
        if (per_bytes > kNoallocScratchBytes)
            throw Error(TACO_ERR_INVALID, "one time's L4 intermediates (" + std::to_string(per_bytes) +
                                              " bytes) exceed the 64 MiB scratch of the allocation-free entry point");
        double2* scr = scratch_on_current_device();
        const long chunk = static_cast<long>(kNoallocScratchBytes / per_bytes);
        const long nf3 = static_cast<long>(nf) * nf * nf;
        const long N4 = N2 * N2;
        L4fArgs F{};
        F.p2f = static_cast<const int*>(pair_to_freq);
        F.ops = static_cast<const double2*>(ops);
        F.N = N;
        F.nf = nf;
        F.num_ops = num_ops;
        F.nf3 = nf3;
        F.scr = scr;
        F.scr_per = per;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        for (long p0 = 0; p0 < static_cast<long>(npos); p0 += chunk) {
            const long np = std::min(chunk, static_cast<long>(npos) - p0);
            F.Fs = static_cast<const double2*>(Fs) + p0 * nf3;
            F.Cs = static_cast<const double2*>(Cs) + p0 * nf3;
            F.Rs = static_cast<const double2*>(Rs) + p0 * nf3;
            F.L4 = static_cast<double2*>(L4_out) + p0 * N4;
            F.GW = GW_out ? static_cast<double2*>(GW_out) + p0 * N4 : nullptr;
            launch_k_l4f(static_cast<unsigned>(np), F.N, lds, st, F);
            HIP_OK(hipGetLastError());
        }
    });
}

int taco_mikx_device(const void* Fs, const void* Cs, const void* Rs, const void* pair_to_freq, int N, int nf,
                     void* M, void* I, void* K, void* X, void* stream) {
    return guarded([&] {
        if (N <= 0 || nf <= 0) throw Error(TACO_ERR_INVALID, "N and nf must be > 0");
        if (!Fs || !Cs || !Rs || !pair_to_freq || !M || !I || !K || !X)
            throw Error(TACO_ERR_INVALID, "device pointers must be non-null");
        const long N2 = static_cast<long>(N) * N, N6 = N2 * N2 * N2;
        if (N2 * 4 > 64 * 1024) throw Error(TACO_ERR_INVALID, "build_mikx_device: N above 128");
        hipLaunchKernelGGL(k_mikx, dim3(static_cast<unsigned>((N6 + kBlock - 1) / kBlock)), dim3(kBlock),
                           static_cast<std::size_t>(N2) * 4, static_cast<hipStream_t>(stream),
                           static_cast<const double2*>(Fs), static_cast<const double2*>(Cs),
                           static_cast<const double2*>(Rs), static_cast<const int*>(pair_to_freq), N, nf,
                           static_cast<double2*>(M), static_cast<double2*>(I), static_cast<double2*>(K),
                           static_cast<double2*>(X));
        HIP_OK(hipGetLastError());
    });
}

int taco_assemble_gw_device(const void* M, const void* I, const void* K, const void* X, const void* ops, int N,
                            int num_ops, void* GW, int raw, void* stream) {
    return guarded([&] {
        if (N <= 0) throw Error(TACO_ERR_INVALID, "N must be > 0");
        if (num_ops <= 0) throw Error(TACO_ERR_INVALID, "assemble_liouvillian: coupling_ops must be non-empty");
        if (!M || !I || !K || !X || !ops || !GW) throw Error(TACO_ERR_INVALID, "device pointers must be non-null");
        const long N2 = static_cast<long>(N) * N, N4 = N2 * N2;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_gw_raw, dim3(static_cast<unsigned>((N4 + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                           static_cast<const double2*>(M), static_cast<const double2*>(I),
                           static_cast<const double2*>(K), static_cast<const double2*>(X),
                           static_cast<const double2*>(ops), N, num_ops, static_cast<double2*>(GW));
        HIP_OK(hipGetLastError());
        if (!raw) {
            const long npairs = N2 * (N2 + 1) / 2;
            hipLaunchKernelGGL(k_gw_sym, dim3(static_cast<unsigned>((npairs + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               st, static_cast<double2*>(GW), static_cast<int>(N2));
            HIP_OK(hipGetLastError());
        }
    });
}

int taco_fcr_batched_dev(taco_ctx* ctx, const void* gamma, const void* omegas, const void* mirror, size_t Nt,
                         int nf, double dt, int i, int j, int k0, int batch, void* F, void* C, void* R) {
    return guarded([&] {
        check_ctx(ctx);
        if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "dt must be > 0");
        if (nf <= 0 || Nt == 0) throw Error(TACO_ERR_INVALID, "empty gamma_series");
        if (batch <= 0) return;
        if (i < 0 || i >= nf || j < 0 || j >= nf || k0 < 0 || k0 + batch > nf)
            throw Error(TACO_ERR_RANGE, "compute_fcr_convolution_batched: (i, j, k0 + batch) outside the buckets");
        if (!gamma || !omegas || !mirror || !F || !C || !R)
            throw Error(TACO_ERR_INVALID, "device pointers must be non-null");
        DeviceGuard g(ctx->device);
        // the pass groups triples by Omega on the host and validates the mirror
        std::vector<double> h_om(static_cast<std::size_t>(nf));
        std::vector<int> h_mi(static_cast<std::size_t>(nf));
        HIP_OK(hipMemcpyAsync(h_om.data(), omegas, h_om.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(h_mi.data(), mirror, h_mi.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        for (int b = 0; b < nf; ++b)
            if (h_mi[static_cast<std::size_t>(b)] < 0 || h_mi[static_cast<std::size_t>(b)] >= nf)
                throw Error(TACO_ERR_INVALID, "mirror index out of range");
        const std::size_t nf3 = static_cast<std::size_t>(nf) * nf * nf;
        std::vector<unsigned char> need(nf3, 0);
        const long tri0 = (static_cast<long>(i) * nf + j) * nf + k0;
        for (int b = 0; b < batch; ++b) need[static_cast<std::size_t>(tri0 + b)] = 7;
        CombineArgs out{};
        out.out_mode = 1;
        out.nf3 = static_cast<long>(nf3);
        out.Fo = static_cast<double2*>(F);
        out.Co = static_cast<double2*>(C);
        out.Ro = static_cast<double2*>(R);
        out.tri0 = tri0;
        FcrPlan pl{static_cast<const double2*>(gamma), Nt, nf, static_cast<const double*>(omegas), h_om.data(),
                   static_cast<const int*>(mirror), dt, Nt, &need};
        run_fcr(ctx, pl, out);
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_tcl2_generator(taco_ctx* ctx, const double* gamma_ab, size_t Nt, int nf, int num_ops, const double* eps,
                        const double* ops, const int* pair_to_freq, int N, const size_t* tidx, size_t ntau,
                        double* L_out) {
    return guarded([&] {
        check_ctx(ctx);
        if (N <= 0 || nf <= 0 || Nt == 0) throw Error(TACO_ERR_INVALID, "N, nf and Nt must be > 0");
        if (num_ops <= 0) throw Error(TACO_ERR_INVALID, "coupling_ops must be non-empty");
        if (!gamma_ab || !eps || !ops || !pair_to_freq || !L_out || (ntau && !tidx))
            throw Error(TACO_ERR_INVALID, "inputs must be non-null");
        const std::size_t N2 = static_cast<std::size_t>(N) * N;
        for (std::size_t e = 0; e < N2; ++e)
            if (pair_to_freq[e] < -1 || pair_to_freq[e] >= nf) throw Error(TACO_ERR_INVALID, "pair_to_freq entry >= nf");
        const std::size_t nt = ntau ? ntau : Nt;
        std::vector<long> tm(nt);
        for (std::size_t q = 0; q < nt; ++q) {
            const std::size_t t = ntau ? tidx[q] : q;
            if (t >= Nt) throw Error(TACO_ERR_RANGE, "time_index out of range");
            tm[q] = static_cast<long>(t);
        }
        // (b, alpha, beta) the reference skips: a zero part, or A_al,b^dag A_be,b == 0
        const std::size_t P = static_cast<std::size_t>(num_ops) * num_ops;
        const std::complex<double>* Ah = reinterpret_cast<const std::complex<double>*>(ops);
        std::vector<unsigned char> skip(static_cast<std::size_t>(nf) * P, 0);
        for (int b = 0; b < nf; ++b)
            for (int be = 0; be < num_ops; ++be)
                for (int al = 0; al < num_ops; ++al) {
                    bool zero = true;
                    for (int r = 0; r < N && zero; ++r)
                        for (int c = 0; c < N && zero; ++c) {
                            std::complex<double> s(0.0, 0.0);
                            for (int q = 0; q < N; ++q)
                                if (pair_to_freq[q + r * N] == b && pair_to_freq[q + c * N] == b)
                                    s += std::conj(Ah[static_cast<std::size_t>(al) * N2 + q + r * N]) *
                                         Ah[static_cast<std::size_t>(be) * N2 + q + c * N];
                            zero = s.real() == 0.0 && s.imag() == 0.0;
                        }
                    skip[(static_cast<std::size_t>(b) * num_ops + be) * num_ops + al] = zero ? 1 : 0;
                }
        DeviceGuard g(ctx->device);
        const std::size_t cols = static_cast<std::size_t>(nf) * P;
        const double2* dg = upload(ctx, ctx->gamma, reinterpret_cast<const double2*>(gamma_ab), Nt * cols);
        const double* de = upload(ctx, ctx->eps, eps, static_cast<std::size_t>(N));
        const double2* dops = upload(ctx, ctx->ops, reinterpret_cast<const double2*>(ops), N2 * num_ops);
        const int* dpf = upload(ctx, ctx->p2f, pair_to_freq, N2);
        ctx->misc.ensure(cols + nt * sizeof(long) + 16);
        unsigned char* dskip = ctx->misc.as<unsigned char>();
        long* dtm = reinterpret_cast<long*>(ctx->misc.as<unsigned char>() + ((cols + 15) / 16) * 16);
        HIP_OK(hipMemcpyAsync(dskip, skip.data(), cols, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(dtm, tm.data(), nt * sizeof(long), hipMemcpyHostToDevice, ctx->stream));
        ctx->l4.ensure(nt * N2 * N2 * 16);
        const long cnt = static_cast<long>(nt * N2 * N2);
        hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, ctx->l4.as<double2>(), cnt);
        Tcl2Args A2{};
        A2.Nt = static_cast<long>(Nt);
        A2.eps = de;
        A2.ops = dops;
        A2.p2f = dpf;
        A2.N = N;
        A2.nf = nf;
        A2.num_ops = num_ops;
        A2.L = ctx->l4.as<double2>();
        A2.order = 2;
        A2.use_lists = 0;
        A2.gamma_ab = dg;
        A2.skip = dskip;
        A2.tmap = dtm;
        launch_tcl2(static_cast<unsigned>(nt), N, tcl2_lds(N), ctx->stream, A2);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(L_out, ctx->l4.p, nt * N2 * N2 * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

int taco_rk4_propagate(taco_ctx* ctx, const double* L_series, const double* L_half, size_t steps, int D, double* r,
                       double dt, size_t sample_every, double* samples_out, size_t* nsamples) {
    return guarded([&] {
        check_ctx(ctx);
        if (!(dt > 0.0)) throw Error(TACO_ERR_INVALID, "propagate_rk4_dense: dt must be > 0");
        if (sample_every == 0) throw Error(TACO_ERR_INVALID, "propagate_rk4_dense: sample_every must be > 0");
        if (steps < 1) throw Error(TACO_ERR_INVALID, "propagate_rk4_dense: L_series must have at least 2 matrices (endpoints)");
        if (D <= 0) throw Error(TACO_ERR_INVALID, "propagate_rk4_dense: L_series and r dimension mismatch");
        DeviceGuard g(ctx->device);
        const std::size_t D2 = static_cast<std::size_t>(D) * D;
        const std::size_t nsam = steps / sample_every + 2;
        const std::size_t bytes = (steps + 1) * D2 * 16 + (L_half ? steps * D2 * 16 : 0) + D * 16 +
                                  (samples_out ? nsam * D * 16 : 0) + 64;
        ctx->rk.ensure(bytes);
        char* base = ctx->rk.as<char>();
        double2* dLs = reinterpret_cast<double2*>(base);
        double2* dLh = L_half ? dLs + (steps + 1) * D2 : nullptr;
        double2* dr = (L_half ? dLh + steps * D2 : dLs + (steps + 1) * D2);
        double2* ds = samples_out ? dr + D : nullptr;
        long* dns = reinterpret_cast<long*>((samples_out ? ds + nsam * D : dr + D));
        HIP_OK(hipMemcpyAsync(dLs, L_series, (steps + 1) * D2 * 16, hipMemcpyHostToDevice, ctx->stream));
        if (L_half) HIP_OK(hipMemcpyAsync(dLh, L_half, steps * D2 * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(dr, r, D * 16, hipMemcpyHostToDevice, ctx->stream));
        rk4_launch(ctx, dLs, dLh, steps, D, dr, dt, sample_every, ds, dns);
        long ns = 0;
        HIP_OK(hipMemcpyAsync(r, dr, D * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipMemcpyAsync(&ns, dns, sizeof(long), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        if (samples_out)
            HIP_OK(hipMemcpy(samples_out, ds, static_cast<std::size_t>(ns) * D * 16, hipMemcpyDeviceToHost));
        if (nsamples) *nsamples = static_cast<std::size_t>(ns);
    });
}

int taco_tcl4_propagate(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas,
                        const int* mirror, const int* pair_to_freq, int N, const double* eps, const double* ops,
                        int num_ops, double dt, size_t steps, int order, double* r, double* samples_out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!(order == 0 || order == 2 || order == 4)) throw Error(TACO_ERR_INVALID, "order must be 0, 2, or 4");
        validate_system(Nt, nf, mirror, pair_to_freq, N, num_ops, dt, TACO_FCR_CONVOLUTION);
        if (steps < 1) throw Error(TACO_ERR_INVALID, "steps must be >= 1");
        if (steps + 1 > Nt) throw Error(TACO_ERR_RANGE, "steps exceed gamma_series rows");
        DeviceGuard g(ctx->device);
        const std::size_t N2 = static_cast<std::size_t>(N) * N, N4 = N2 * N2;
        const double2* dg = upload(ctx, ctx->gamma, reinterpret_cast<const double2*>(gamma), Nt * nf);
        const double* dom = upload(ctx, ctx->omegas, omegas, static_cast<std::size_t>(nf));
        const int* dmi = upload(ctx, ctx->mirror, mirror, static_cast<std::size_t>(nf));
        const int* dpf = upload(ctx, ctx->p2f, pair_to_freq, N2);
        const double2* dops = upload(ctx, ctx->ops, reinterpret_cast<const double2*>(ops), N2 * num_ops);
        const double* deps = upload(ctx, ctx->eps, eps, static_cast<std::size_t>(N));
        const std::size_t T = steps + 1;
        if (order == 4) {
            std::vector<std::size_t> all(T);
            std::iota(all.begin(), all.end(), std::size_t{0});
            l4_core(ctx, dg, Nt, nf, dom, omegas, dmi, dpf, pair_to_freq, N, dops, num_ops, dt, all.data(), T,
                    nullptr);
        } else {
            ctx->l4.ensure(T * N4 * 16);
            HIP_OK(hipMemsetAsync(ctx->l4.p, 0, T * N4 * 16, ctx->stream));
        }
        Tcl2Args A2{};
        A2.gamma = dg;
        A2.Nt = static_cast<long>(Nt);
        A2.eps = deps;
        A2.ops = dops;
        A2.p2f = dpf;
        A2.N = N;
        A2.use_lists = tcl2_lists_enabled() ? 1 : 0;
        A2.nf = nf;
        A2.num_ops = num_ops;
        A2.L = ctx->l4.as<double2>();
        A2.order = order;
        launch_tcl2(static_cast<unsigned>(T), A2.N, tcl2_lds(A2.N), ctx->stream, A2);
        HIP_OK(hipGetLastError());
        ctx->misc.ensure((T + 1) * N2 * 16 + 64);
        double2* dr = ctx->misc.as<double2>();
        double2* ds = dr + N2;
        HIP_OK(hipMemcpyAsync(dr, r, N2 * 16, hipMemcpyHostToDevice, ctx->stream));
        rk4_launch(ctx, ctx->l4.as<double2>(), nullptr, steps, static_cast<int>(N2), dr, dt, 1,
                   samples_out ? ds : nullptr, nullptr);
        HIP_OK(hipMemcpyAsync(r, dr, N2 * 16, hipMemcpyDeviceToHost, ctx->stream));
        if (samples_out) HIP_OK(hipMemcpyAsync(samples_out, ds, T * N2 * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

struct taco_batch {
    taco_ctx* ctx = nullptr;
    taco_engine::BatchRun run;
};

int taco_batch_create(taco_ctx* ctx, const taco_batch_system* systems, int nsys, taco_batch** out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw Error(TACO_ERR_INVALID, "null output handle");
        *out = nullptr;
        if (nsys < 0 || (nsys > 0 && !systems)) throw Error(TACO_ERR_INVALID, "invalid system list");
        DeviceGuard g(ctx->device);
        auto b = std::make_unique<taco_batch>();
        b->ctx = ctx;
        BatchRun& B = b->run;
        std::vector<Rk4Sys> rk;
        for (int q = 0; q < nsys; ++q) {
            const taco_batch_system& d = systems[q];
            validate_system(d.Nt, d.nf, d.mirror, d.pair_to_freq, d.N, d.num_ops, d.dt, TACO_FCR_CONVOLUTION);
            if (!d.gamma || !d.omegas || !d.mirror || !d.pair_to_freq || !d.ops)
                throw Error(TACO_ERR_INVALID, "system inputs must be non-null");
            if (!(d.order == 0 || d.order == 2 || d.order == 4)) throw Error(TACO_ERR_INVALID, "order must be 0, 2, or 4");
            if (d.ntau && !d.tidx) throw Error(TACO_ERR_INVALID, "null time index list");
            if (d.steps && (!d.r0 || !d.eps)) throw Error(TACO_ERR_INVALID, "RK4 needs r0 and eps");
            if (d.steps + 1 > d.Nt && d.steps) throw Error(TACO_ERR_RANGE, "steps exceed gamma_series rows");
            auto S = std::make_unique<BatchSys>();
            S->Nt = d.Nt;
            S->nf = d.nf;
            S->N = d.N;
            S->num_ops = d.num_ops;
            S->dt = d.dt;
            S->ntau = d.ntau;
            S->steps = d.steps;
            S->order = d.order;
            const std::size_t N2 = static_cast<std::size_t>(d.N) * d.N, N4 = N2 * N2;
            auto put = [&](DevBuf& buf, const void* host, std::size_t bytes) {
                buf.ensure(std::max<std::size_t>(bytes, 16));
                if (bytes) HIP_OK(hipMemcpy(buf.p, host, bytes, hipMemcpyHostToDevice));
            };
            put(S->gamma, d.gamma, d.Nt * d.nf * 16);
            put(S->omegas, d.omegas, d.nf * sizeof(double));
            put(S->mirror, d.mirror, d.nf * sizeof(int));
            put(S->p2f, d.pair_to_freq, N2 * sizeof(int));
            put(S->ops, d.ops, N2 * d.num_ops * 16);
            if (d.steps) {
                put(S->eps, d.eps, d.N * sizeof(double));
                put(S->r0, d.r0, N2 * 16);
            }
            // L4 times: the requested ones and, for a TCL4 run, every RK4 step.
            std::vector<std::size_t> want(d.tidx, d.tidx + d.ntau);
            if (d.steps && d.order == 4)
                for (std::size_t t = 0; t <= d.steps; ++t) want.push_back(t);
            if (!want.empty()) {
                const TimePlan tp = plan_times(want.data(), want.size(), d.Nt);
                S->npos = static_cast<long>(tp.uniq.size());
                S->T = tp.T;
                put(S->t2pos, tp.t2pos.data(), tp.T * sizeof(int));
                std::vector<int> rows(tp.req2u.begin(), tp.req2u.begin() + d.ntau);
                put(S->rows_out, rows.data(), rows.size() * sizeof(int));
                if (d.steps && d.order == 4) {
                    std::vector<int> gen(tp.req2u.begin() + d.ntau, tp.req2u.end());
                    put(S->rows_gen, gen.data(), gen.size() * sizeof(int));
                }
                const std::vector<unsigned char> need = needed_triples(d.pair_to_freq, d.N, d.nf);
                FcrPlan pl{S->gamma.as<double2>(), d.Nt, d.nf, S->omegas.as<double>(), d.omegas, S->mirror.as<int>(),
                           d.dt, tp.T, &need};
                S->fcr = fcr_prepare(ctx, pl, FcrLists{&S->jobtri, &S->jobmask, &S->clsstart, &S->jobrow, &S->phtab});
                const std::size_t nf3 = static_cast<std::size_t>(d.nf) * d.nf * d.nf;
                if (static_cast<std::size_t>(S->npos) > (~std::size_t{0} / 48) / nf3)
                    throw Error(TACO_ERR_OVERFLOW, "FCR buffer too large");
                ctx->slices.ensure(3 * nf3 * static_cast<std::size_t>(S->npos) * 16);
                ctx->l4.ensure(N4 * static_cast<std::size_t>(S->npos) * 16);
                ctx->misc.ensure(static_cast<std::size_t>(l4f_chunk(d.N, d.nf, d.num_ops, S->npos) *
                                                          l4f_per_pos(d.N, d.nf, d.num_ops)) * 16);
                const std::size_t lds = l4_lds_bytes(d.N, d.num_ops);
                if (lds <= 160 * 1024) set_l4_lds(ctx, lds);  // for TACO_L4_KERNEL=lds
            }
            S->l4out.ensure(std::max<std::size_t>(d.ntau * N4 * 16, 16));
            if (d.steps) {
                S->lgen.ensure((d.steps + 1) * N4 * 16);
                S->rho.ensure((d.steps + 2) * N2 * 16);
                Rk4Sys r{};
                r.Ls = S->lgen.as<double2>();
                r.r0 = S->r0.as<double2>();
                r.samples = S->rho.as<double2>();
                r.r = r.samples + (d.steps + 1) * N2;
                r.steps = static_cast<long>(d.steps);
                r.dt = d.dt;
                r.D = static_cast<int>(N2);
                const bool rows = rk4_rows_threads(r.D) > 0;
                r.nstage = rows ? 0 : rk4_nstage(r.D, 6, static_cast<long>(d.steps) + 1);
                rk.push_back(r);
                B.rk_lds = std::max(B.rk_lds, rk4_lds(r.D, 6, r.nstage));
                B.rk_block = std::max(B.rk_block, rows ? rk4_rows_threads(r.D) : rk4_block(r.D));
                B.rk_bytes += static_cast<double>(d.steps + 1) * r.D * (r.D + 1) * 16.0;
            }
            B.sys.push_back(std::move(S));
        }
        B.nrk = static_cast<long>(rk.size());
        if (B.nrk) {
            if (B.rk_lds > kRkLds) throw Error(TACO_ERR_INVALID, "RK4 kernel supports D <= 1706");
            B.rkargs.ensure(rk.size() * sizeof(Rk4Sys));
            HIP_OK(hipMemcpy(B.rkargs.p, rk.data(), rk.size() * sizeof(Rk4Sys), hipMemcpyHostToDevice));
        }
        HIP_OK(hipStreamSynchronize(ctx->stream));
        *out = b.release();
    });
}

int taco_batch_run(taco_batch* b, int use_graph) {
    return guarded([&] {
        if (!b) throw Error(TACO_ERR_INVALID, "null taco_batch");
        taco_ctx* c = b->ctx;
        BatchRun& B = b->run;
        DeviceGuard g(c->device);
        if (!use_graph || !B.ran_eager) {
            // The first run is eager: it picks the FFT tile configurations
            // (autotune synchronises) that a captured graph then keeps.
            batch_enqueue(c, B);
            HIP_OK(hipStreamSynchronize(c->stream));
            B.ran_eager = true;
            if (!use_graph) return;
        }
        // The graph holds the addresses of the context's workspace; a call
        // that has reallocated any device buffer since the capture makes it
        // stale, so it is captured again.
        if (B.exec && B.gen != g_dev_realloc.load()) B.drop_graph();
        if (!B.exec) {
            const bool prof = c->prof.on;
            c->prof.on = false;
            HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            try {
                batch_enqueue(c, B);
            } catch (...) {
                hipGraph_t junk = nullptr;
                (void)hipStreamEndCapture(c->stream, &junk);
                if (junk) (void)hipGraphDestroy(junk);
                c->prof.on = prof;
                throw;
            }
            c->prof.on = prof;
            HIP_OK(hipStreamEndCapture(c->stream, &B.graph));
            HIP_OK(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
            B.gen = g_dev_realloc.load();
        }
        HIP_OK(hipGraphLaunch(B.exec, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int taco_batch_l4(taco_batch* b, int s, double* L4_out) {
    return guarded([&] {
        if (!b) throw Error(TACO_ERR_INVALID, "null taco_batch");
        if (s < 0 || s >= static_cast<int>(b->run.sys.size())) throw Error(TACO_ERR_RANGE, "system index out of range");
        const BatchSys& S = *b->run.sys[static_cast<std::size_t>(s)];
        const std::size_t N2 = static_cast<std::size_t>(S.N) * S.N;
        DeviceGuard g(b->ctx->device);
        if (S.ntau) HIP_OK(hipMemcpy(L4_out, S.l4out.p, S.ntau * N2 * N2 * 16, hipMemcpyDeviceToHost));
    });
}

int taco_batch_rho(taco_batch* b, int s, double* samples_out, double* r_out) {
    return guarded([&] {
        if (!b) throw Error(TACO_ERR_INVALID, "null taco_batch");
        if (s < 0 || s >= static_cast<int>(b->run.sys.size())) throw Error(TACO_ERR_RANGE, "system index out of range");
        const BatchSys& S = *b->run.sys[static_cast<std::size_t>(s)];
        if (!S.steps) throw Error(TACO_ERR_INVALID, "system has no RK4 steps");
        const std::size_t N2 = static_cast<std::size_t>(S.N) * S.N;
        DeviceGuard g(b->ctx->device);
        if (samples_out) HIP_OK(hipMemcpy(samples_out, S.rho.p, (S.steps + 1) * N2 * 16, hipMemcpyDeviceToHost));
        if (r_out)
            HIP_OK(hipMemcpy(r_out, S.rho.as<double2>() + (S.steps + 1) * N2, N2 * 16, hipMemcpyDeviceToHost));
    });
}

int taco_batch_destroy(taco_batch* b) {
    return guarded([&] {
        if (!b) return;
        DeviceGuard g(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        delete b;
    });
}

int taco_batched_run(taco_ctx* ctx, const taco_batch_system* systems, int nsys, double* const* L4_out,
                     double* const* rho_out) {
    taco_batch* b = nullptr;
    int rc = taco_batch_create(ctx, systems, nsys, &b);
    if (rc) return rc;
    rc = taco_batch_run(b, 0);
    for (int s = 0; rc == TACO_OK && s < nsys; ++s) {
        if (L4_out && L4_out[s]) rc = taco_batch_l4(b, s, L4_out[s]);
        if (rc == TACO_OK && rho_out && rho_out[s] && systems[s].steps) rc = taco_batch_rho(b, s, rho_out[s], nullptr);
    }
    const std::string msg = rc ? std::string(taco_last_error()) : std::string();
    const int rc2 = taco_batch_destroy(b);
    if (rc) set_last_error(msg.c_str());  // keep the failure's message past the clean-up
    return rc ? rc : rc2;
}

int taco_tape_forward(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                      const int* pair_to_freq, int N, const double* eps, const double* ops, int num_ops, double dt,
                      size_t steps, int order, const double* r0, size_t checkpoint_every, double* samples_out,
                      taco_tape** out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw Error(TACO_ERR_INVALID, "null output handle");
        *out = nullptr;
        DeviceGuard g(ctx->device);
        const taco_tape_set in{gamma, Nt, nf, omegas, mirror, pair_to_freq, N, eps, ops, num_ops, dt, steps, order, r0};
        auto tp = tape_make(ctx, in, checkpoint_every);
        tape_forward(*tp, r0, samples_out);
        *out = tp.release();
    });
}

int taco_mtape_forward(taco_ctx* ctx, const taco_tape_set* sets, int nsets, size_t checkpoint_every,
                       double* const* samples_out, taco_mtape** out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw Error(TACO_ERR_INVALID, "null output handle");
        *out = nullptr;
        if (!sets || nsets < 1) throw Error(TACO_ERR_INVALID, "need at least one parameter set");
        for (int s = 1; s < nsets; ++s)
            if (sets[s].N != sets[0].N || sets[s].steps != sets[0].steps)
                throw Error(TACO_ERR_INVALID, "parameter sets of a multi-set tape must share N and steps");
        DeviceGuard g(ctx->device);
        auto mt = std::make_unique<taco_mtape>();
        mt->ctx = ctx;
        mtape_claim(*mt);
        for (int s = 0; s < nsets; ++s) mt->sets.push_back(tape_make(ctx, sets[s], checkpoint_every));
        mtape_forward(*mt, sets, samples_out);
        *out = mt.release();
    });
}

int taco_mtape_adjoint(taco_mtape* mt, const double* const* samples_bar, double* const* gamma_bar,
                       double* const* omega_bar, double* const* ops_bar, double* const* eps_bar,
                       double* const* r0_bar) {
    return guarded([&] {
        if (!mt) throw Error(TACO_ERR_INVALID, "null taco_mtape");
        if (!samples_bar) throw Error(TACO_ERR_INVALID, "null samples_bar");
        for (std::size_t s = 0; s < mt->sets.size(); ++s)
            if (!samples_bar[s]) throw Error(TACO_ERR_INVALID, "null samples_bar entry");
        DeviceGuard g(mt->ctx->device);
        mtape_adjoint(*mt, samples_bar, gamma_bar, omega_bar, ops_bar, eps_bar, r0_bar);
    });
}

int taco_mtape_forward_bath(taco_ctx* ctx, const taco_tape_set* sets, const taco_bath_params* baths, int nsets,
                            size_t checkpoint_every, double* const* samples_out, taco_mtape** out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw Error(TACO_ERR_INVALID, "null output handle");
        *out = nullptr;
        if (!sets || !baths || nsets < 1) throw Error(TACO_ERR_INVALID, "need at least one parameter set and bath");
        for (int s = 1; s < nsets; ++s)
            if (sets[s].N != sets[0].N || sets[s].steps != sets[0].steps)
                throw Error(TACO_ERR_INVALID, "parameter sets of a multi-set tape must share N and steps");
        DeviceGuard g(ctx->device);
        auto mt = std::make_unique<taco_mtape>();
        mt->ctx = ctx;
        mtape_claim(*mt);
        for (int s = 0; s < nsets; ++s) {
            const BathSrc bath{baths[s].kind, baths[s].p0, baths[s].p1, baths[s].beta};
            mt->sets.push_back(tape_make(ctx, sets[s], checkpoint_every, &bath));
        }
        mtape_forward(*mt, sets, samples_out);
        *out = mt.release();
    });
}

int taco_mtape_adjoint_bath(taco_mtape* mt, const double* const* samples_bar, double* const* bath_bar,
                            double* const* omega_bar, double* const* ops_bar, double* const* eps_bar,
                            double* const* r0_bar) {
    return guarded([&] {
        if (!mt) throw Error(TACO_ERR_INVALID, "null taco_mtape");
        if (!samples_bar || !bath_bar) throw Error(TACO_ERR_INVALID, "null samples_bar or bath_bar");
        for (std::size_t s = 0; s < mt->sets.size(); ++s)
            if (!samples_bar[s] || !bath_bar[s]) throw Error(TACO_ERR_INVALID, "null samples_bar or bath_bar entry");
        DeviceGuard g(mt->ctx->device);
        mtape_adjoint(*mt, samples_bar, nullptr, omega_bar, ops_bar, eps_bar, r0_bar, bath_bar);
    });
}

int taco_mtape_keeps_slices(const taco_mtape* mt) { return mt && mt->keep ? 1 : 0; }

int taco_mtape_set_slices(const taco_mtape* mt, size_t s, size_t* bytes) {
    if (!mt || s >= mt->sets.size()) return -1;
    const int mode = mt->keep_set[s];
    if (bytes) {
        const taco_tape& tp = *mt->sets[s];
        const auto& n = mt->pk_n[s];
        *bytes = mode == 1 ? tape_slice_bytes(tp)
                 : mode == 2 ? static_cast<std::size_t>(n[0] + n[1] + n[2]) * tp.T * 16
                             : 0;
    }
    return mode;
}

int taco_mtape_destroy(taco_mtape* mt) {
    return guarded([&] {
        if (!mt) return;
        DeviceGuard g(mt->ctx->device);
        (void)hipStreamSynchronize(mt->ctx->stream);
        for (auto& tp : mt->sets) tape_pool_return(*tp);
        delete mt;
    });
}

int taco_tape_adjoint(taco_tape* tape, const double* samples_bar, double* gamma_bar, double* omega_bar,
                      double* ops_bar, double* eps_bar, double* r0_bar) {
    return guarded([&] {
        if (!tape) throw Error(TACO_ERR_INVALID, "null taco_tape");
        if (!samples_bar) throw Error(TACO_ERR_INVALID, "null samples_bar");
        DeviceGuard g(tape->ctx->device);
        tape_adjoint(*tape, samples_bar, gamma_bar, omega_bar, ops_bar, eps_bar, r0_bar);
    });
}

int taco_tape_forward_bath(taco_ctx* ctx, int kind, double p0, double p1, double beta, size_t Nt, int nf,
                           const double* omegas, const int* mirror, const int* pair_to_freq, int N, const double* eps,
                           const double* ops, int num_ops, double dt, size_t steps, int order, const double* r0,
                           size_t checkpoint_every, double* samples_out, taco_tape** out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw Error(TACO_ERR_INVALID, "null output handle");
        *out = nullptr;
        DeviceGuard g(ctx->device);
        const taco_tape_set in{nullptr, Nt, nf, omegas, mirror, pair_to_freq, N, eps, ops, num_ops, dt, steps, order, r0};
        const BathSrc bath{kind, p0, p1, beta};
        auto tp = tape_make(ctx, in, checkpoint_every, &bath);
        tape_forward(*tp, r0, samples_out);
        *out = tp.release();
    });
}

int taco_tape_adjoint_bath(taco_tape* tape, const double* samples_bar, double* bath_bar, double* omega_bar,
                           double* ops_bar, double* eps_bar, double* r0_bar) {
    return guarded([&] {
        if (!tape) throw Error(TACO_ERR_INVALID, "null taco_tape");
        if (!samples_bar || !bath_bar) throw Error(TACO_ERR_INVALID, "null samples_bar or bath_bar");
        DeviceGuard g(tape->ctx->device);
        tape_adjoint(*tape, samples_bar, nullptr, omega_bar, ops_bar, eps_bar, r0_bar, bath_bar);
    });
}

int taco_tape_destroy(taco_tape* tape) {
    return guarded([&] {
        if (!tape) return;
        DeviceGuard g(tape->ctx->device);
        (void)hipStreamSynchronize(tape->ctx->stream);
        tape_pool_return(*tape);
        delete tape;
    });
}

int taco_adjoint(taco_ctx* ctx, const double* gamma, size_t Nt, int nf, const double* omegas, const int* mirror,
                 const int* pair_to_freq, int N, const double* eps, const double* ops, int num_ops, double dt,
                 size_t steps, int order, const double* r0, size_t checkpoint_every, const double* samples_bar,
                 double* gamma_bar, double* omega_bar, double* ops_bar, double* eps_bar, double* r0_bar) {
    taco_tape* tp = nullptr;
    int rc = taco_tape_forward(ctx, gamma, Nt, nf, omegas, mirror, pair_to_freq, N, eps, ops, num_ops, dt, steps,
                               order, r0, checkpoint_every, nullptr, &tp);
    if (rc) return rc;
    rc = taco_tape_adjoint(tp, samples_bar, gamma_bar, omega_bar, ops_bar, eps_bar, r0_bar);
    const std::string msg = rc ? std::string(taco_last_error()) : std::string();
    const int rc2 = taco_tape_destroy(tp);
    if (rc) set_last_error(msg.c_str());
    return rc ? rc : rc2;
}

}  // extern "C"
