// Row-group RK4 kernels bodies for D = N^2 with N = 2..8 (rk4_dense.cpp:133-160,
// 246-320 for the step; the reverse walk has no reference counterpart).
//
// Included by engine.hip inside namespace taco_engine, after the complex
// helpers.  G consecutive lanes own one output element of every
// matrix-vector product: lane g sums the terms k = g, g + G, g + 2G, ... in
// order with fused multiply-adds, and the G partial sums are combined by a
// butterfly over the group (DPP lane swaps; a + b == b + a exactly, so every
// lane of a group ends with the same value).  Generators never pass through
// LDS: each lane keeps its CK entries of the two generators of the current
// step in registers, and the loads of later generators are issued up to
// S - 1 steps ahead (a ring of S register slots), which hides global latency
// behind the sequential steps.  Vectors go through small LDS buffers with
// one barrier per RK4 stage.  Forward, batched forward and segment
// recompute share rk4_rows, so their states are identical bit for bit.
#pragma once

template <int DT>
struct RowGroup {
    static constexpr int G = DT <= 16 ? 4 : 8;   // lanes per output element
    static constexpr int R = 64 / G;             // output elements per wave
    static constexpr int W = (DT + R - 1) / R;   // waves
    static constexpr int CK = (DT + G - 1) / G;  // terms per lane
    static constexpr int S = DT <= 16 ? 4 : 3;   // generator ring slots (forward)
    static constexpr int SA = 3;                 // ring slots of the reverse walk
};

// Threads of a row-group workgroup for D (0 when D has no row-group path).
__host__ __device__ constexpr int rk4_rows_threads(int D) {
    return D == 4 ? 64 * RowGroup<4>::W
         : D == 9 ? 64 * RowGroup<9>::W
         : D == 16 ? 64 * RowGroup<16>::W
         : D == 25 ? 64 * RowGroup<25>::W
         : D == 36 ? 64 * RowGroup<36>::W
         : D == 49 ? 64 * RowGroup<49>::W
         : D == 64 ? 64 * RowGroup<64>::W
         : 0;
}

// One 64-bit value from another lane of the same row of 16 (DPP control CTRL).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(b)), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(b >> 32)), CTRL, 0xF, 0xF,
                                               false);
    const unsigned long long r = (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                 static_cast<unsigned>(lo);
    return __longlong_as_double(static_cast<long long>(r));
}

// Sum over the G consecutive lanes of a group: xor 1, xor 2 (quad
// permutations), then the half-row mirror (lane i <-> 7 - i) for G = 8.
template <int G>
__device__ __forceinline__ double2 group_sum(double2 v) {
    static_assert(G == 4 || G == 8, "row groups of 4 or 8 lanes");
    v.x = v.x + dpp_f64<0xB1>(v.x);
    v.y = v.y + dpp_f64<0xB1>(v.y);
    v.x = v.x + dpp_f64<0x4E>(v.x);
    v.y = v.y + dpp_f64<0x4E>(v.y);
    if constexpr (G == 8) {
        v.x = v.x + dpp_f64<0x141>(v.x);
        v.y = v.y + dpp_f64<0x141>(v.y);
    }
    return v;
}

// s + a b and s + conj(a) b with fused multiply-adds.
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 s) {
    s.x = fma(a.x, b.x, s.x);
    s.x = fma(-a.y, b.y, s.x);
    s.y = fma(a.x, b.y, s.y);
    s.y = fma(a.y, b.x, s.y);
    return s;
}
__device__ __forceinline__ double2 cfma_conj(double2 a, double2 b, double2 s) {
    s.x = fma(a.x, b.x, s.x);
    s.x = fma(a.y, b.y, s.x);
    s.y = fma(a.x, b.y, s.y);
    s.y = fma(-a.y, b.x, s.y);
    return s;
}

// Lane's CK entries of a D x D column-major matrix M.  ROWS: lane owns row
// `o` and columns g + k G (forward products); otherwise column `o` and rows
// g + k G (adjoint products).  Entries outside the matrix read M[0]: a finite
// value that only reaches sums of groups or terms no result keeps.  The
// loads are unconditional and unmasked, so nothing waits for them before
// their values are used, steps later.
template <int DT, bool ROWS>
__device__ __forceinline__ void load_gen(double2 (&v)[RowGroup<DT>::CK], const double2* __restrict__ M, int o,
                                         int g) {
    constexpr int G = RowGroup<DT>::G;
#pragma unroll
    for (int k = 0; k < RowGroup<DT>::CK; ++k) {
        const int q = g + k * G;
        const bool ok = o < DT && q < DT;
        v[k] = M[ok ? (ROWS ? o + q * DT : q + o * DT) : 0];
    }
}

// The group's product element: sum_k m[k] * v[g + k G] (CONJ: conj(m[k])).
template <int DT, bool CONJ>
__device__ __forceinline__ double2 group_dot(const double2 (&m)[RowGroup<DT>::CK], const double2* v, int g) {
    constexpr int G = RowGroup<DT>::G;
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < RowGroup<DT>::CK; ++k) {
        const int q = g + k * G;
        if (DT % G == 0 || q < DT) s = CONJ ? cfma_conj(m[k], v[q], s) : cfma(m[k], v[q], s);
    }
    return group_sum<G>(s);
}

// One RK4 step of rk4_rows.  L0 / L1: the lane's entries of generators n and
// n + 1; L0's slot is refilled with generator n + S once the midpoint is
// formed.
template <int DT, bool STASH>
struct RowsForward {
    using RG = RowGroup<DT>;
    static constexpr int CK = RG::CK;
    static constexpr long D2 = static_cast<long>(DT) * DT;
    const double2* __restrict__ Ls;
    double2 *X, *B1, *B2, *samples, *seg;
    long n1, every, ns, cnt;
    double dt, hdt, s6;
    int r, g;
    bool own;
    double2 xr;

    __device__ __forceinline__ void step(long n, double2 (&L0)[CK], const double2 (&L1)[CK]) {
        if (samples) {
            if (cnt == 0) {
                if (own) samples[ns * DT + r] = xr;
                ++ns;
                cnt = every;
            }
            --cnt;
        }
        double2* row = STASH ? seg + n * 4 * DT : nullptr;
        if constexpr (STASH) {
            if (own) row[r] = xr;
        }
        // k1 = L_n x
        const double2 k1 = group_dot<DT, false>(L0, X, g);
        double2 Lh[CK];
#pragma unroll
        for (int k = 0; k < CK; ++k) Lh[k] = cscale(cadd(L0[k], L1[k]), 0.5);
        double2 u = cadd(xr, cscale(k1, hdt));
        if (own) B1[r] = u;
        if constexpr (STASH) {
            if (own) row[DT + r] = u;
        }
        double2 acc = k1;
        __syncthreads();
        // k2 = Lh u1
        const double2 k2 = group_dot<DT, false>(Lh, B1, g);
        u = cadd(xr, cscale(k2, hdt));
        if (own) B2[r] = u;
        if constexpr (STASH) {
            if (own) row[2 * DT + r] = u;
        }
        acc = cadd(acc, cscale(k2, 2.0));
        __syncthreads();
        // k3 = Lh u2
        const double2 k3 = group_dot<DT, false>(Lh, B2, g);
        u = cadd(xr, cscale(k3, dt));
        if (own) B1[r] = u;
        if constexpr (STASH) {
            if (own) row[3 * DT + r] = u;
        }
        acc = cadd(acc, cscale(k3, 2.0));
        __syncthreads();
        // k4 = L_{n+1} u3
        const double2 k4 = group_dot<DT, false>(L1, B1, g);
        acc = cadd(acc, k4);
        xr = cadd(xr, cscale(acc, s6));
        if (own) X[r] = xr;
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        load_gen<DT, true>(L0, Ls + min(n + RG::S, n1) * D2, r, g);
    }
};

// RK4 steps n0 .. n1-1 on x (x_in -> x_out, if given), step n using Ls[n],
// Ls[n + 1] and the midpoint 0.5 (Ls[n] + Ls[n + 1]).  samples (optional): x
// before every step with (n - n0) % every == 0 and after the last one.
// STASH: the state r_n and stage points u1, u2, u3 of step n go to
// seg[n][4][D].  LDS: 3 D values.  Whole blocks of S steps run in a loop with
// no early exit (so each ring slot keeps its registers across iterations);
// the remaining steps follow it.
template <int DT, bool STASH>
__device__ __forceinline__ void rk4_rows(double2* sh, const double2* __restrict__ Ls, long n0, long n1,
                                         const double2* x_in, double2* x_out, double dt, long every,
                                         double2* samples, long* nsamples, double2* seg) {
    using RG = RowGroup<DT>;
    constexpr int CK = RG::CK, S = RG::S;
    constexpr long D2 = static_cast<long>(DT) * DT;
    const int lane = static_cast<int>(threadIdx.x & 63);
    RowsForward<DT, STASH> F;
    F.Ls = Ls;
    F.X = sh;
    F.B1 = sh + DT;
    F.B2 = sh + 2 * DT;
    F.samples = samples;
    F.seg = seg;
    F.n1 = n1;
    F.every = every;
    F.ns = 0;
    F.cnt = 0;
    F.dt = dt;
    F.hdt = 0.5 * dt;
    F.s6 = dt / 6.0;
    F.g = lane % RG::G;
    F.r = (RG::W == 1 ? 0 : static_cast<int>(threadIdx.x >> 6) * RG::R) + lane / RG::G;
    F.own = F.r < DT && F.g == 0;
    F.xr = make_double2(0.0, 0.0);
    if (F.r < DT) F.xr = x_in[F.r];
    if (F.own) F.X[F.r] = F.xr;
    // ring slot (m - n0) % S holds generator m
    double2 Lr[S][CK];
#pragma unroll
    for (int q = 0; q < S; ++q) load_gen<DT, true>(Lr[q], Ls + min(n0 + q, n1) * D2, F.r, F.g);
    __syncthreads();
    long nb = n0;
    for (; nb + S <= n1; nb += S) {
#pragma unroll
        for (int j = 0; j < S; ++j) F.step(nb + j, Lr[j], Lr[(j + 1) % S]);
    }
#pragma unroll
    for (int j = 0; j + 1 < S; ++j)
        if (nb + j < n1) F.step(nb + j, Lr[j], Lr[(j + 1) % S]);
    if (samples) {
        if (F.own) samples[F.ns * DT + F.r] = F.xr;
        ++F.ns;
    }
    if (F.own && x_out) x_out[F.r] = F.xr;
    if (threadIdx.x == 0 && nsamples) *nsamples = F.ns;
}

// State, stage points and gbar of step n at element c (element 0 for c
// outside D, whose results are never stored).
template <int DT, class Args>
__device__ __forceinline__ void load_points(double2 (&p)[4], double2& gb, const Args& A, long n, int c) {
    const int cc = c < DT ? c : 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) p[v] = A.seg[n * 4 * DT + v * DT + cc];
    gb = A.gbar[n * DT + cc];
}

// Reverse walk of rk4_rows over all steps (one wave: D <= 16).  Group c owns
// element c of every cotangent vector and column c of the generator
// cotangents.  Per step n, with lam the cotangent of r_{n+1}:
//   kb4 = s lam, ub3 = L1^H kb4, kb3 = 2 s lam + dt ub3,
//   ub2 = Lh^H kb3, kb2 = 2 s lam + h ub2, ub1 = Lh^H kb2,
//   kb1 = s lam + h ub1, ub0 = L0^H kb1,
//   lam_n = lam + ub3 + ub2 + ub1 + ub0 + gbar_n,
//   Lbar_{n+1} += kb4 u3^H + (kb3 u2^H + kb2 u1^H) / 2,
//   Lbar_n     += kb1 r_n^H + (kb3 u2^H + kb2 u1^H) / 2,
// the L0 part carried in registers to the next (earlier) step.  Generators,
// stage points and gbar are prefetched S - 1 steps ahead.  LDS: 5 D values.
template <int DT, class Args, bool DEFER = false>
struct RowsAdjoint {
    using RG = RowGroup<DT>;
    static constexpr int CK = RG::CK, S = RG::SA, G = RG::G;
    static constexpr long D2 = static_cast<long>(DT) * DT;
    const Args* A;
    double2 *K4, *K3, *K2, *K1;
    double h, s6;
    int c, g, par;
    bool col, own;
    double2 lam;
    double2 Lc[CK];

    // Step n.  L1 / L0: the lane's entries of generators n + 1 and n; Sg / Gb:
    // r_n, u1, u2, u3 and gbar_n at element c.  L1's slot is refilled with
    // generator n + 1 - S and the points with those of step n - S.
    __device__ __forceinline__ void step(long n, double2 (&L1)[CK], const double2 (&L0)[CK], double2 (&Sg)[4],
                                         double2& Gb) {
        const double2* K4p = K4 + par * DT;
        const double2 kb4 = cscale(lam, s6);  // = K4p[c]
        // ub3 = L1^H kb4
        const double2 ub3 = group_dot<DT, true>(L1, K4p, g);
        const double2 kb3 = cadd(cscale(lam, 2.0 * s6), cscale(ub3, A->dt));
        double2 rb = cadd(lam, ub3);
        if (own) K3[c] = kb3;
        double2 Lh[CK];
#pragma unroll
        for (int k = 0; k < CK; ++k) Lh[k] = cscale(cadd(L0[k], L1[k]), 0.5);
        __builtin_amdgcn_sched_barrier(0);
        load_gen<DT, false>(L1, A->Ls + max(n + 1 - S, 0L) * D2, c, g);
        __syncthreads();
        // ub2 = Lh^H kb3
        const double2 ub2 = group_dot<DT, true>(Lh, K3, g);
        const double2 kb2 = cadd(cscale(lam, 2.0 * s6), cscale(ub2, h));
        rb = cadd(rb, ub2);
        if (own) K2[c] = kb2;
        __syncthreads();
        // ub1 = Lh^H kb2
        const double2 ub1 = group_dot<DT, true>(Lh, K2, g);
        const double2 kb1 = cadd(cscale(lam, s6), cscale(ub1, h));
        rb = cadd(rb, ub1);
        if (own) K1[c] = kb1;
        __syncthreads();
        // ub0 = L0^H kb1
        const double2 ub0 = group_dot<DT, true>(L0, K1, g);
        if constexpr (DEFER) {
            // the step's kb1..kb4 for k_rk4_lbar, which forms every Lbar(t)
            // afterwards with these sums (all times in parallel)
            if (own) {
                double2* v = A->vec + n * 4 * DT;
                v[c] = kb1;
                v[DT + c] = kb2;
                v[2 * DT + c] = kb3;
                v[3 * DT + c] = kb4;
            }
        } else {
            // generator cotangents of column c, rows g + k G
            const double2 xc = Sg[0], u1c = Sg[1], u2c = Sg[2], u3c = Sg[3];
            double2* Lb1 = A->Lbar + (n + 1) * D2;
#pragma unroll
            for (int k = 0; k < CK; ++k) {
                const int rr = g + k * G;
                if (col && rr < DT) {
                    const double2 lh = cadd(cmul(K3[rr], cconj(u2c)), cmul(K2[rr], cconj(u1c)));
                    Lb1[rr + c * DT] = cadd(Lc[k], cadd(cmul(K4p[rr], cconj(u3c)), cscale(lh, 0.5)));
                    Lc[k] = cadd(cmul(K1[rr], cconj(xc)), cscale(lh, 0.5));
                }
            }
        }
        lam = cadd(cadd(rb, ub0), Gb);
        par ^= 1;
        if (own) K4[par * DT + c] = cscale(lam, s6);
        __builtin_amdgcn_sched_barrier(0);
        load_points<DT>(Sg, Gb, *A, max(n - S, 0L), c);
        __syncthreads();
    }
};

// DEFER: the walk stores each step's kb1..kb4 (A.vec) instead of forming
// Lbar; k_rk4_lbar then forms Lbar(t) from them and the stage points, with
// the walk's own sums (the same values; off the walk's sequential path).
template <int DT, class Args, bool DEFER = false>
__device__ __forceinline__ void rk4_rows_adjoint(double2* sh, const Args& A) {
    using RG = RowGroup<DT>;
    static_assert(RG::W == 1, "one wave");
    constexpr int CK = RG::CK, S = RG::SA, G = RG::G;
    constexpr long D2 = static_cast<long>(DT) * DT;
    const int lane = static_cast<int>(threadIdx.x & 63);
    RowsAdjoint<DT, Args, DEFER> B;
    B.A = &A;
    B.K4 = sh;  // two buffers, alternating by step
    B.K3 = sh + 2 * DT;
    B.K2 = sh + 3 * DT;
    B.K1 = sh + 4 * DT;
    B.h = 0.5 * A.dt;
    B.s6 = A.dt / 6.0;
    B.g = lane % G;
    B.c = lane / G;
    B.col = B.c < DT;
    B.own = B.col && B.g == 0;
    B.par = 0;
    const long steps = A.steps;
    // lam first: its load is then the oldest one in flight when the loop
    // starts, so no wait for the ring loads lands in the loop
    B.lam = A.gbar[steps * DT + (B.col ? B.c : 0)];
    // ring slot (steps - m) % S holds generator m; slot (steps - 1 - n) % S
    // the stage points and gbar of step n
    double2 Lr[S][CK], Sg[S][4], Gb[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        load_gen<DT, false>(Lr[q], A.Ls + max(steps - q, 0L) * D2, B.c, B.g);
        load_points<DT>(Sg[q], Gb[q], A, max(steps - 1 - q, 0L), B.c);
    }
#pragma unroll
    for (int k = 0; k < CK; ++k) B.Lc[k] = make_double2(0.0, 0.0);
    if (B.own) B.K4[B.c] = cscale(B.lam, B.s6);
    __syncthreads();
    // whole blocks of S steps (no early exit), then the rest
    long nb = steps - 1;
    for (; nb + 1 >= S; nb -= S) {
#pragma unroll
        for (int j = 0; j < S; ++j) B.step(nb - j, Lr[j], Lr[(j + 1) % S], Sg[j], Gb[j]);
    }
#pragma unroll
    for (int j = 0; j + 1 < S; ++j)
        if (nb - j >= 0) B.step(nb - j, Lr[j], Lr[(j + 1) % S], Sg[j], Gb[j]);
    if constexpr (!DEFER) {
#pragma unroll
        for (int k = 0; k < CK; ++k) {
            const int rr = B.g + k * G;
            if (B.col && rr < DT) A.Lbar[rr + B.c * DT] = B.Lc[k];
        }
    }
    if (B.own) A.r0bar[B.c] = B.lam;
}
