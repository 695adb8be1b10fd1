// Reverse-mode kernels of the TCL4 pipeline: RK4 -> fused L4 -> F/C/R.
//
// Included once by engine.hip after the forward kernels (it uses their device
// helpers).  The reference has no derivative code; these are the hand-written
// adjoints of the forward passes above.  Convention for every complex
// cotangent z_bar: dLoss = Re sum conj(z_bar) dz, i.e. z_bar = dLoss/dRe z +
// i dLoss/dIm z (torch's convention for real losses).
#pragma once

// y = L^H x (column r of L conjugated, column stride ld), summed over c in order.
// (General D: column blocks of kMvBlk loads issued ahead, as in matvec_rows.)
template <int DT>
__device__ __forceinline__ void matvec_h(const double2* __restrict__ L, int ld, const double2* x, double2* y, int D) {
    const int n = DT ? DT : D;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double2 s = make_double2(0.0, 0.0);
        const double2* col = L + static_cast<long>(r) * ld;
        if constexpr (DT == 0) {
            int c = 0;
            for (; c + kMvBlk <= n; c += kMvBlk) {
                double2 v[kMvBlk];
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) v[j] = col[c + j];
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) s = cadd(s, cmul(cconj(v[j]), x[c + j]));
            }
            for (; c < n; ++c) s = cadd(s, cmul(cconj(col[c]), x[c]));
        } else {
#pragma unroll
            for (int c = 0; c < DT; ++c) s = cadd(s, cmul(cconj(col[c]), x[c]));
        }
        y[r] = s;
    }
}

// y = (0.5 (L0 + L1))^H x.
template <int DT>
__device__ __forceinline__ void matvec_h_half(const double2* __restrict__ L0, const double2* __restrict__ L1, int ld,
                                              const double2* x, double2* y, int D) {
    const int n = DT ? DT : D;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double2 s = make_double2(0.0, 0.0);
        const long o = static_cast<long>(r) * ld;
        if constexpr (DT == 0) {
            int c = 0;
            for (; c + kMvBlk <= n; c += kMvBlk) {
                double2 v0[kMvBlk], v1[kMvBlk];
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) {
                    v0[j] = L0[o + c + j];
                    v1[j] = L1[o + c + j];
                }
#pragma unroll
                for (int j = 0; j < kMvBlk; ++j) s = cadd(s, cmul(cconj(cscale(cadd(v0[j], v1[j]), 0.5)), x[c + j]));
            }
            for (; c < n; ++c) s = cadd(s, cmul(cconj(cscale(cadd(L0[o + c], L1[o + c]), 0.5)), x[c]));
        } else {
#pragma unroll
            for (int c = 0; c < DT; ++c)
                s = cadd(s, cmul(cconj(cscale(cadd(L0[o + c], L1[o + c]), 0.5)), x[c]));
        }
        y[r] = s;
    }
}

// ---------------------------------------------------------------------------
// RK4 adjoint.  Forward step n (rk4_step): k1 = L0 r, u1 = r + h k1,
// k2 = Lh u1, u2 = r + h k2, k3 = Lh u2, u3 = r + dt k3, k4 = L1 u3,
// r' = r + s (k1 + 2 k2 + 2 k3 + k4), h = dt/2, s = dt/6, Lh = (L0 + L1)/2.
// Two launches.  k_rk4_recompute: one workgroup per checkpoint segment (K
// steps) recomputes, from the checkpoint ck[q] = r_{qK} and with the forward
// step's own arithmetic, the states r_n and stage points u1..u3 of its
// segment; segments are independent, so all run at once.  k_rk4_adjoint:
// one workgroup walks n = steps-1 .. 0, mapping lam_{n+1} to
// lam_n = M_n^H lam_{n+1} + gbar_n and forming the cotangents of L_n and
// L_{n+1}.  LDS: vectors of D, then (nstage >= 2) a window of generators
// staged with column stride D + 1.  Lbar(t) collects the L1 part of step t-1
// and the L0 part of step t: the latter is carried in registers
// (D^2 <= 16 x 256) and each Lbar(t) stored once.  Larger D (generators read
// from global memory): the walk reads the conjugate transposes LsH (written
// beforehand by k_conj_transpose, so every product is a row-coalesced
// matrix-vector product with the same terms in the same order), stores only
// the step's cotangent vectors kb1..kb4, and k_rk4_lbar forms every Lbar(t)
// afterwards, all times at once, with the walk's own sums.
// ---------------------------------------------------------------------------
struct Rk4AdjArgs {
    const double2* Ls;    // [steps + 1][D * D] generators
    const double2* ck;    // [ceil(steps / K)][D] checkpoint states r_{qK}
    const double2* gbar;  // [steps + 1][D] dLoss/dr_n
    double2* Lbar;        // [steps + 1][D * D]
    double2* seg;         // [steps][4][D] state r_n and stage points u1, u2, u3 of step n
    double2* r0bar;       // [D]
    long steps;
    int D;
    int K;
    double dt;
    int nstage;
    const double2* LsH = nullptr;  // [steps + 1][D * D] conjugate transposes (walk without staging)
    const double2* LhH = nullptr;  // [steps][D * D] their midpoints (null: formed on the fly)
    double2* vec = nullptr;        // [steps][4][D] kb1..kb4 of every step (Lbar formed by k_rk4_lbar)
};

constexpr int kRkCarry = 16;

// Makes generators L_a..L_b resident in the LDS window M (column stride
// D + 1), restaging the window [first, last] when they are not.
struct GenWindow {
    long lo = 0, hi = -1;
    __device__ __forceinline__ void need(double2* M, const double2* Ls, long a, long b, long first, long last, int D) {
        if (a >= lo && b <= hi) return;
        __syncthreads();
        lo = first;
        hi = last;
        stage_mats(M, Ls, lo, static_cast<int>(hi - lo + 1), D);
        __syncthreads();
    }
};

template <int DT, bool STAGED>
__device__ __forceinline__ void rk4_recompute_body(double2* sh, const Rk4AdjArgs& A, long sg) {
    const int D = A.D;
    double2* x = sh;
    double2* k1 = x + D;
    double2* k2 = k1 + D;
    double2* k3 = k2 + D;
    double2* k4 = k3 + D;
    double2* tmp = k4 + D;
    double2* M = tmp + D;
    const long D2 = static_cast<long>(D) * D;
    const int ld = STAGED ? D + 1 : D;
    GenWindow w;
    const long s0 = sg * A.K, s1 = min(s0 + A.K, A.steps);
    for (int e = threadIdx.x; e < D; e += blockDim.x) x[e] = A.ck[sg * D + e];
    __syncthreads();
    for (long n = s0; n < s1; ++n) {
        double2* row = A.seg + n * 4 * D;
        for (int e = threadIdx.x; e < D; e += blockDim.x) row[e] = x[e];
        if constexpr (STAGED) {
            w.need(M, A.Ls, n, n + 1, n, min(n + A.nstage - 1, s1), D);
            rk4_step<DT>(M + (n - w.lo) * D * ld, M + (n + 1 - w.lo) * D * ld, ld, nullptr, x, k1, k2, k3, k4, tmp,
                         D, A.dt, row + D);
        } else {
            rk4_step<DT>(A.Ls + n * D2, A.Ls + (n + 1) * D2, ld, nullptr, x, k1, k2, k3, k4, tmp, D, A.dt, row + D);
        }
    }
}

// Row-group recompute (rk4_rows with the stage points stashed) for the sizes
// that have one: the same arithmetic as the forward run, so the states equal
// the forward's bit for bit.  One workgroup per checkpoint segment.
template <int DT>
__global__ __launch_bounds__(64 * RowGroup<DT>::W) void k_rk4_recompute_rows(Rk4AdjArgs A) {
    extern __shared__ double2 sh[];
    const long sg = blockIdx.x;
    const long s0 = sg * A.K, s1 = min(s0 + A.K, A.steps);
    rk4_rows<DT, true>(sh, A.Ls, s0, s1, A.ck + sg * DT, nullptr, A.dt, 1, nullptr, nullptr, A.seg);
}

__global__ __launch_bounds__(kBlock) void k_rk4_recompute(Rk4AdjArgs A) {
    extern __shared__ double2 sh[];
    const long sg = blockIdx.x;
    if (A.nstage >= 2) {
#define TACO_RK4_CALL(DT) rk4_recompute_body<DT, true>(sh, A, sg)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    } else {
#define TACO_RK4_CALL(DT) rk4_recompute_body<DT, false>(sh, A, sg)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    }
}

template <int DT, bool STAGED>
__device__ __forceinline__ void rk4_adjoint_body(double2* sh, const Rk4AdjArgs& A) {
    const int D = A.D;
    double2* x = sh;          // r_n
    double2* k3 = x + D;      // cotangent of k3
    double2* k4 = k3 + D;     // cotangent of k4
    double2* tmp = k4 + D;
    double2* lam = tmp + D;   // cotangent of r_{n+1}, then of r_n
    double2* u1 = lam + D;
    double2* u2 = u1 + D;
    double2* u3 = u2 + D;
    double2* kb1 = u3 + D;    // cotangents of k1, k2
    double2* kb2 = kb1 + D;
    double2* rb = kb2 + D;    // cotangent of r_n being accumulated
    double2* M = rb + D;      // staged generators
    const long D2 = static_cast<long>(D) * D;
    const int ld = STAGED ? D + 1 : D;
    const bool carry = D2 <= static_cast<long>(kRkCarry) * blockDim.x;
    double2 Lc[kRkCarry];     // L0 part of the step after the current one (carry mode)
#pragma unroll
    for (int q = 0; q < kRkCarry; ++q) Lc[q] = make_double2(0.0, 0.0);
    GenWindow w;
    const double h = 0.5 * A.dt, s6 = A.dt / 6.0;
    for (int e = threadIdx.x; e < D; e += blockDim.x) lam[e] = A.gbar[A.steps * D + e];
    __syncthreads();
    for (long n = A.steps - 1; n >= 0; --n) {
        const double2* L0;
        const double2* L1;
        const double2* H0 = nullptr;  // conjugate transposes (LsH path)
        const double2* H1 = nullptr;
        if constexpr (STAGED) {
            w.need(M, A.Ls, n, n + 1, max(0L, n + 2 - A.nstage), n + 1, D);
            L0 = M + (n - w.lo) * D * ld;
            L1 = L0 + D * ld;
        } else {
            L0 = A.Ls + n * D2;
            L1 = L0 + D2;
            if (A.LsH) {
                H0 = A.LsH + n * D2;
                H1 = H0 + D2;
            }
        }
        const double2* row = A.seg + n * 4 * D;
        for (int e = threadIdx.x; e < D; e += blockDim.x) {
            x[e] = row[e];
            u1[e] = row[D + e];
            u2[e] = row[2 * D + e];
            u3[e] = row[3 * D + e];
            k4[e] = cscale(lam[e], s6);                                                    // kb4
        }
        __syncthreads();
        // reverse: r' = r + s (k1 + 2 k2 + 2 k3 + k4)
        if (H1) matvec_rows<DT>(H1, D, k4, tmp, D); else matvec_h<DT>(L1, ld, k4, tmp, D);  // ub3
        __syncthreads();
        for (int e = threadIdx.x; e < D; e += blockDim.x) {
            k3[e] = cadd(cscale(lam[e], 2.0 * s6), cscale(tmp[e], A.dt));                  // kb3
            rb[e] = cadd(lam[e], tmp[e]);
        }
        __syncthreads();
        if (A.LhH && H0) matvec_rows<DT>(A.LhH + n * D2, D, k3, tmp, D);                             // ub2
        else if (H0) matvec_half<DT>(H0, H1, D, k3, tmp, D);
        else matvec_h_half<DT>(L0, L1, ld, k3, tmp, D);
        __syncthreads();
        for (int e = threadIdx.x; e < D; e += blockDim.x) {
            kb2[e] = cadd(cscale(lam[e], 2.0 * s6), cscale(tmp[e], h));
            rb[e] = cadd(rb[e], tmp[e]);
        }
        __syncthreads();
        if (A.LhH && H0) matvec_rows<DT>(A.LhH + n * D2, D, kb2, tmp, D);                            // ub1
        else if (H0) matvec_half<DT>(H0, H1, D, kb2, tmp, D);
        else matvec_h_half<DT>(L0, L1, ld, kb2, tmp, D);
        __syncthreads();
        for (int e = threadIdx.x; e < D; e += blockDim.x) {
            kb1[e] = cadd(cscale(lam[e], s6), cscale(tmp[e], h));
            rb[e] = cadd(rb[e], tmp[e]);
        }
        __syncthreads();
        if (H0) matvec_rows<DT>(H0, D, kb1, tmp, D); else matvec_h<DT>(L0, ld, kb1, tmp, D);
        __syncthreads();
        // generator cotangents: Lbar_h = kb3 u2^H + kb2 u1^H, split half/half
        double2* Lb0 = A.Lbar + n * D2;
        double2* Lb1 = A.Lbar + (n + 1) * D2;
        if (!carry && A.vec) {
            double2* v = A.vec + n * 4 * D;
            for (int e = threadIdx.x; e < D; e += blockDim.x) {
                v[e] = kb1[e];
                v[D + e] = kb2[e];
                v[2 * D + e] = k3[e];
                v[3 * D + e] = k4[e];
            }
        } else if (carry) {
#pragma unroll
            for (int q = 0; q < kRkCarry; ++q) {
                const long e = threadIdx.x + static_cast<long>(q) * blockDim.x;
                if (e < D2) {
                    const int r = static_cast<int>(e % D), c = static_cast<int>(e / D);
                    const double2 lh = cadd(cmul(k3[r], cconj(u2[c])), cmul(kb2[r], cconj(u1[c])));
                    Lb1[e] = cadd(Lc[q], cadd(cmul(k4[r], cconj(u3[c])), cscale(lh, 0.5)));
                    Lc[q] = cadd(cmul(kb1[r], cconj(x[c])), cscale(lh, 0.5));
                }
            }
        } else {
            // in place, kU elements per thread at a time with their loads issued
            // first (the same expressions as the remainder loop)
            constexpr int kU = 8;
            const long stride = blockDim.x;
            long e0 = threadIdx.x;
            for (; e0 + (kU - 1) * stride < D2; e0 += kU * stride) {
                double2 a0[kU], a1[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    a0[u] = Lb0[e0 + u * stride];
                    a1[u] = Lb1[e0 + u * stride];
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const long e = e0 + u * stride;
                    const int r = static_cast<int>(e % D), c = static_cast<int>(e / D);
                    const double2 lh = cadd(cmul(k3[r], cconj(u2[c])), cmul(kb2[r], cconj(u1[c])));
                    Lb0[e] = cadd(a0[u], cadd(cmul(kb1[r], cconj(x[c])), cscale(lh, 0.5)));
                    Lb1[e] = cadd(a1[u], cadd(cmul(k4[r], cconj(u3[c])), cscale(lh, 0.5)));
                }
            }
            for (long e = e0; e < D2; e += stride) {
                const int r = static_cast<int>(e % D), c = static_cast<int>(e / D);
                const double2 lh = cadd(cmul(k3[r], cconj(u2[c])), cmul(kb2[r], cconj(u1[c])));
                Lb0[e] = cadd(Lb0[e], cadd(cmul(kb1[r], cconj(x[c])), cscale(lh, 0.5)));
                Lb1[e] = cadd(Lb1[e], cadd(cmul(k4[r], cconj(u3[c])), cscale(lh, 0.5)));
            }
        }
        for (int e = threadIdx.x; e < D; e += blockDim.x) lam[e] = cadd(cadd(rb[e], tmp[e]), A.gbar[n * D + e]);
        __syncthreads();
    }
    if (carry) {
#pragma unroll
        for (int q = 0; q < kRkCarry; ++q) {
            const long e = threadIdx.x + static_cast<long>(q) * blockDim.x;
            if (e < D2) A.Lbar[e] = Lc[q];
        }
    }
    for (int e = threadIdx.x; e < D; e += blockDim.x) A.r0bar[e] = lam[e];
}

// Row-group reverse walk for D <= 16 (one wave of 64 threads).
template <int DT>
__global__ __launch_bounds__(64) void k_rk4_adjoint_rows(Rk4AdjArgs A) {
    extern __shared__ double2 sh[];
    if (A.vec)
        rk4_rows_adjoint<DT, Rk4AdjArgs, true>(sh, A);
    else
        rk4_rows_adjoint<DT>(sh, A);
}

__global__ __launch_bounds__(kBlock) void k_rk4_adjoint(Rk4AdjArgs A) {
    extern __shared__ double2 sh[];
    if (A.nstage >= 2) {
#define TACO_RK4_CALL(DT) rk4_adjoint_body<DT, true>(sh, A)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    } else {
#define TACO_RK4_CALL(DT) rk4_adjoint_body<DT, false>(sh, A)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    }
}

// dst[t] = src[t]^H for count D x D matrices (tiles of 16 x 16 through LDS,
// coalesced reads and writes).  Grid: (tiles of one matrix, count).
__global__ __launch_bounds__(kBlock) void k_conj_transpose(const double2* __restrict__ src, double2* __restrict__ dst,
                                                           int D, long count) {
    __shared__ double2 tile[16][17];
    const int tiles = (D + 15) / 16;
    const int ti = static_cast<int>(blockIdx.x) % tiles, tj = static_cast<int>(blockIdx.x) / tiles;
    const long m = blockIdx.y;
    if (m >= count) return;
    const long D2 = static_cast<long>(D) * D;
    const int lr = threadIdx.x % 16, lc = threadIdx.x / 16;
    const int r = ti * 16 + lr, c = tj * 16 + lc;  // source element (r, c)
    if (r < D && c < D) tile[lc][lr] = src[m * D2 + r + static_cast<long>(c) * D];
    __syncthreads();
    const int r2 = tj * 16 + lr, c2 = ti * 16 + lc;  // destination element (r2, c2) = conj(src(c2, r2))
    if (r2 < D && c2 < D) dst[m * D2 + r2 + static_cast<long>(c2) * D] = cconj(tile[lr][lc]);
}

// Lbar(t), t = 0..steps, from the walk's vectors: the L0 part of step t
// (kb1 x^H + (kb3 u2^H + kb2 u1^H)/2) added to zero, then the L1 part of
// step t-1 (kb4 u3^H + (kb3 u2^H + kb2 u1^H)/2): the sums the in-place walk
// forms, in its order.  Grid: (element blocks of D^2, steps + 1[, tapes]).
__device__ __forceinline__ void rk4_lbar_body(const Rk4AdjArgs& A) {
    const int D = A.D;
    const long D2 = static_cast<long>(D) * D;
    const long e = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    const long t = blockIdx.y;
    if (e >= D2 || t > A.steps) return;
    const int r = static_cast<int>(e % D), c = static_cast<int>(e / D);
    double2 acc = make_double2(0.0, 0.0);
    if (t < A.steps) {
        const double2* v = A.vec + t * 4 * D;
        const double2* sg = A.seg + t * 4 * D;
        const double2 lh = cadd(cmul(v[2 * D + r], cconj(sg[2 * D + c])), cmul(v[D + r], cconj(sg[D + c])));
        acc = cadd(acc, cadd(cmul(v[r], cconj(sg[c])), cscale(lh, 0.5)));
    }
    if (t >= 1) {
        const double2* v = A.vec + (t - 1) * 4 * D;
        const double2* sg = A.seg + (t - 1) * 4 * D;
        const double2 lh = cadd(cmul(v[2 * D + r], cconj(sg[2 * D + c])), cmul(v[D + r], cconj(sg[D + c])));
        acc = cadd(acc, cadd(cmul(v[3 * D + r], cconj(sg[3 * D + c])), cscale(lh, 0.5)));
    }
    A.Lbar[t * D2 + e] = acc;
}

__global__ __launch_bounds__(kBlock) void k_rk4_lbar(Rk4AdjArgs A) { rk4_lbar_body(A); }

__global__ __launch_bounds__(kBlock) void k_rk4_lbar_multi(const Rk4AdjArgs* As) {
    const Rk4AdjArgs A = As[blockIdx.z];
    rk4_lbar_body(A);
}

// Many tapes of one D, steps and K in one launch: the recompute with
// blockIdx.y choosing the tape, the walk with one workgroup per tape; each
// workgroup runs the single-tape body on its own arguments.
template <int DT>
__global__ __launch_bounds__(64 * RowGroup<DT>::W) void k_rk4_recompute_rows_multi(const Rk4AdjArgs* As) {
    extern __shared__ double2 sh[];
    const Rk4AdjArgs A = As[blockIdx.y];
    const long sg = blockIdx.x;
    const long s0 = sg * A.K, s1 = min(s0 + A.K, A.steps);
    rk4_rows<DT, true>(sh, A.Ls, s0, s1, A.ck + sg * DT, nullptr, A.dt, 1, nullptr, nullptr, A.seg);
}

__global__ __launch_bounds__(kBlock) void k_rk4_recompute_multi(const Rk4AdjArgs* As) {
    extern __shared__ double2 sh[];
    const Rk4AdjArgs A = As[blockIdx.y];
    const long sg = blockIdx.x;
    if (A.nstage >= 2) {
#define TACO_RK4_CALL(DT) rk4_recompute_body<DT, true>(sh, A, sg)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    } else {
#define TACO_RK4_CALL(DT) rk4_recompute_body<DT, false>(sh, A, sg)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    }
}

template <int DT>
__global__ __launch_bounds__(64) void k_rk4_adjoint_rows_multi(const Rk4AdjArgs* As) {
    extern __shared__ double2 sh[];
    const Rk4AdjArgs A = As[blockIdx.x];
    if (A.vec)
        rk4_rows_adjoint<DT, Rk4AdjArgs, true>(sh, A);
    else
        rk4_rows_adjoint<DT>(sh, A);
}

__global__ __launch_bounds__(kBlock) void k_rk4_adjoint_multi(const Rk4AdjArgs* As) {
    extern __shared__ double2 sh[];
    const Rk4AdjArgs A = As[blockIdx.x];
    if (A.nstage >= 2) {
#define TACO_RK4_CALL(DT) rk4_adjoint_body<DT, true>(sh, A)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    } else {
#define TACO_RK4_CALL(DT) rk4_adjoint_body<DT, false>(sh, A)
        TACO_RK4_DISPATCH(A.D, TACO_RK4_CALL)
#undef TACO_RK4_CALL
    }
}

// ---------------------------------------------------------------------------
// Fused L4 adjoint, one workgroup per time t.  From Lbar(t) (the RK4 layout of
// L4(n,m;i,j) = GW(n,i;m,j)) it forms Tbar = GWbar + GWbar^H (GW = T + T^H)
// and, for every NAKZWAN term sigma * e1 e2 e3 e4 * S of T(n,i;m,j), adds
// sigma Tbar conj(e1 e2 e3 e4) to the slice entries S gathers from and
// sigma Tbar conj(S * prod of the other three) to each operator entry.  Slice
// cotangents accumulate in LDS (or global when they do not fit) with atomic
// adds; operator cotangents in LDS, written per time.
// ---------------------------------------------------------------------------
struct L4AdjArgs {
    const double2* Lbar;  // [T][N^4]
    const double2* Fs;    // [T][nf3] forward slices
    const double2* Cs;
    const double2* Rs;
    const int* p2f;
    const double2* ops;
    int N, nf, num_ops;
    long nf3;
    double2* sbar;        // [T][3][nf3]; zero on entry unless lds_slices
    double2* Abar;        // [T][num_ops][N^2]
    int lds_slices;
};

__device__ __forceinline__ void atomic_cadd(double2* p, double2 v) {
    atomicAdd(&p->x, v.x);
    atomicAdd(&p->y, v.y);
}

__global__ __launch_bounds__(kBlock) void k_l4_adjoint(L4AdjArgs A) {
    extern __shared__ double2 sm[];
    const int N = A.N, N2 = N * N, N4 = N2 * N2;
    const long t = blockIdx.x;
    double2* Tb = sm;                                   // N^4
    double2* opsS = Tb + N4;                            // num_ops N^2
    double2* Ab = opsS + A.num_ops * N2;                // num_ops N^2
    double2* sb = Ab + A.num_ops * N2;                  // 3 nf3 (lds_slices)
    int* f = reinterpret_cast<int*>(sb + (A.lds_slices ? 3 * A.nf3 : 0));
    const double2* Lb = A.Lbar + t * N4;
    for (int idx = threadIdx.x; idx < N4; idx += kBlock) {
        const int ru = idx % N2, cu = idx / N2;
        const int n = ru % N, i = ru / N, m = cu % N, j = cu / N;
        const double2 g1 = Lb[(n + m * N) + (i + j * N) * N2];
        const double2 g2 = Lb[(m + n * N) + (j + i * N) * N2];
        Tb[idx] = cadd(g1, cconj(g2));
    }
    for (int e = threadIdx.x; e < A.num_ops * N2; e += kBlock) {
        opsS[e] = A.ops[e];
        Ab[e] = make_double2(0.0, 0.0);
    }
    if (A.lds_slices)
        for (long e = threadIdx.x; e < 3 * A.nf3; e += kBlock) sb[e] = make_double2(0.0, 0.0);
    for (int e = threadIdx.x; e < N2; e += kBlock) f[e] = A.p2f[e];
    __syncthreads();
    const double2* Fp = A.Fs + t * A.nf3;
    const double2* Cp = A.Cs + t * A.nf3;
    const double2* Rp = A.Rs + t * A.nf3;
    double2* sF = A.lds_slices ? sb : A.sbar + t * 3 * A.nf3;
    double2* sC = sF + A.nf3;
    double2* sR = sC + A.nf3;
    const long nf = A.nf;
    auto id3 = [nf](int x, int y, int z) { return (static_cast<long>(x) * nf + y) * nf + z; };
    auto F2 = [f, N](int r, int c) { return f[r + c * N]; };
    // One term sigma * e1 e2 e3 e4 * S with S = sum of up to two slice
    // entries (w1 * slice1[x1] + w2 * slice2[x2]); d1..d4 are the operator
    // cotangent slots of e1..e4.
    auto term = [&](double sg, double2 Tv, double2 e1, double2 e2, double2 e3, double2 e4, double2* d1, double2* d2,
                    double2* d3, double2* d4, double2 S, double2* s1, long x1, double w1, double2* s2, long x2,
                    double w2) {
        const double2 p12 = cmul(e1, e2), p34 = cmul(e3, e4);
        const double2 Sb = cscale(cmul(Tv, cconj(cmul(p12, p34))), sg);
        atomic_cadd(s1 + x1, cscale(Sb, w1));
        if (s2) atomic_cadd(s2 + x2, cscale(Sb, w2));
        const double2 TS = cscale(cmul(Tv, cconj(S)), sg);
        atomic_cadd(d1, cmul(TS, cconj(cmul(e2, p34))));
        atomic_cadd(d2, cmul(TS, cconj(cmul(e1, p34))));
        atomic_cadd(d3, cmul(TS, cconj(cmul(p12, e4))));
        atomic_cadd(d4, cmul(TS, cconj(cmul(p12, e3))));
    };
    for (int idx = threadIdx.x; idx < N4; idx += kBlock) {
        const int row_u = idx % N2, col_u = idx / N2;
        const int n = row_u % N, i = row_u / N, m = col_u % N, j = col_u / N;
        const double2 Tv = Tb[idx];
        if (Tv.x == 0.0 && Tv.y == 0.0) continue;
        for (int iA = 0; iA < A.num_ops; ++iA) {
            const double2* Ap = opsS + iA * N2;
            double2* AbA = Ab + iA * N2;
            for (int iB = 0; iB < A.num_ops; ++iB) {
                const double2* Bp = opsS + iB * N2;
                double2* AbB = Ab + iB * N2;
#define AA(r, c) Ap[(r) + (c) * N]
#define BB(r, c) Bp[(r) + (c) * N]
#define DA(r, c) (AbA + (r) + (c) * N)
#define DB(r, c) (AbB + (r) + (c) * N)
                for (int a = 0; a < N; ++a) {
                    for (int b = 0; b < N; ++b) {
                        const long xM1 = id3(F2(b, a), F2(b, i), F2(n, b)), xM2 = id3(F2(b, a), F2(n, a), F2(a, i));
                        const long xI = id3(F2(a, n), F2(b, i), F2(n, b));
                        const long xK = id3(F2(i, b), F2(n, a), F2(a, i));
                        const long x1 = id3(F2(a, n), F2(j, b), F2(n, i));
                        const long x2 = id3(F2(i, a), F2(j, b), F2(n, i));
                        const long x3 = id3(F2(b, a), F2(j, m), F2(a, i));
                        const long x4 = id3(F2(i, b), F2(j, m), F2(a, i));
                        const double2 Mv = csub(Fp[xM1], Rp[xM2]);
                        const double2 Iv = Fp[xI], Kv = Rp[xK];
                        const double2 X1 = cadd(Cp[x1], Rp[x1]), X2 = cadd(Cp[x2], Rp[x2]);
                        const double2 X3 = cadd(Cp[x3], Rp[x3]), X4 = cadd(Cp[x4], Rp[x4]);
                        // T = -(t1 - t2 + t3 + t4 - t5 - t6 + t7)
                        term(-1.0, Tv, BB(n, a), AA(a, b), BB(b, i), AA(j, m), DB(n, a), DA(a, b), DB(b, i), DA(j, m),
                             Mv, sF, xM1, 1.0, sR, xM2, -1.0);
                        term(+1.0, Tv, AA(n, a), BB(a, b), BB(b, i), AA(j, m), DA(n, a), DB(a, b), DB(b, i), DA(j, m),
                             Iv, sF, xI, 1.0, nullptr, 0, 0.0);
                        term(-1.0, Tv, BB(n, a), BB(a, b), AA(b, i), AA(j, m), DB(n, a), DB(a, b), DA(b, i), DA(j, m),
                             Kv, sR, xK, 1.0, nullptr, 0, 0.0);
                        term(-1.0, Tv, AA(n, a), BB(a, i), BB(j, b), AA(b, m), DA(n, a), DB(a, i), DB(j, b), DA(b, m),
                             X1, sC, x1, 1.0, sR, x1, 1.0);
                        term(+1.0, Tv, BB(n, a), AA(a, i), BB(j, b), AA(b, m), DB(n, a), DA(a, i), DB(j, b), DA(b, m),
                             X2, sC, x2, 1.0, sR, x2, 1.0);
                        term(+1.0, Tv, AA(n, a), AA(a, b), BB(b, i), BB(j, m), DA(n, a), DA(a, b), DB(b, i), DB(j, m),
                             X3, sC, x3, 1.0, sR, x3, 1.0);
                        term(-1.0, Tv, AA(n, a), BB(a, b), AA(b, i), BB(j, m), DA(n, a), DB(a, b), DA(b, i), DB(j, m),
                             X4, sC, x4, 1.0, sR, x4, 1.0);
                    }
                }
                if (j == m) {
                    for (int a = 0; a < N; ++a)
                        for (int b = 0; b < N; ++b)
                            for (int c = 0; c < N; ++c) {
                                const long yM1 = id3(F2(c, b), F2(c, i), F2(a, c));
                                const long yM2 = id3(F2(c, b), F2(a, b), F2(b, i));
                                const long yK = id3(F2(i, c), F2(a, b), F2(b, i));
                                const long yI = id3(F2(b, a), F2(c, i), F2(a, c));
                                const double2 Mv = csub(Fp[yM1], Rp[yM2]);
                                const double2 Kv = Rp[yK], Iv = Fp[yI];
                                term(+1.0, Tv, AA(n, a), BB(a, b), AA(b, c), BB(c, i), DA(n, a), DB(a, b), DA(b, c),
                                     DB(c, i), Mv, sF, yM1, 1.0, sR, yM2, -1.0);
                                term(+1.0, Tv, AA(n, a), BB(a, b), BB(b, c), AA(c, i), DA(n, a), DB(a, b), DB(b, c),
                                     DA(c, i), Kv, sR, yK, 1.0, nullptr, 0, 0.0);
                                term(-1.0, Tv, AA(n, a), AA(a, b), BB(b, c), BB(c, i), DA(n, a), DA(a, b), DB(b, c),
                                     DB(c, i), Iv, sF, yI, 1.0, nullptr, 0, 0.0);
                            }
                }
#undef AA
#undef BB
#undef DA
#undef DB
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < A.num_ops * N2; e += kBlock) A.Abar[t * A.num_ops * N2 + e] = Ab[e];
    if (A.lds_slices)
        for (long e = threadIdx.x; e < 3 * A.nf3; e += kBlock) A.sbar[t * 3 * A.nf3 + e] = sb[e];
}

// ---------------------------------------------------------------------------
// Factorised L4 adjoint: the reverse of k_l4f, O(N^5 + N^3 nf) per time
// instead of the O(N^6) term walk above.  Per operator pair (A, B) the
// forward builds
//   Q(a,i,g)  = sum_b A(a,b) B(b,i) X(f(b,a),g,f(a,i)) - B(a,b) A(b,i) X(f(i,b),g,f(a,i)),
//   S123(n,i) = sum_{a,b} -B(n,a)A(a,b)B(b,i) M + A(n,a)B(a,b)B(b,i) I - B(n,a)B(a,b)A(b,i) K,
//   S67(n,i,g) = sum_a A(n,a) Q(a,i,g),  D(n,i) = -sum_a A(n,a) S123(a,i),
//   T(n,i;m,j) = A(j,m) S123(n,i) + B(j,m) S67(n,i,f(j,m))
//                - sum_b B(j,b) A(b,m) Q(n,i,f(j,b)) + d_jm D(n,i),
// with X = C + R, M = F(f(b,a),f(b,i),f(n,b)) - R(f(b,a),f(n,a),f(a,i)),
// I = F(f(a,n),f(b,i),f(n,b)), K = R(f(i,b),f(n,a),f(a,i)) (the E
// intermediate of k_l4f equals -S123, and its Wd equals Q).  The reverse
// walks these back: Tbar -> (S123bar, S67bar, Dbar, U = A^H-contracted
// Tbar) -> Qbar, S123bar -> slice and operator cotangents.  Every cotangent
// entry has one owner thread per stage that sums its terms in a fixed order
// (gathers over host-built term lists for the slices), so gradients are
// reproducible bit for bit.  Slice cotangents sit in LDS when 3 nf^3 fits,
// else in global memory.  Each block walks times blockIdx.x, blockIdx.x +
// gridDim.x, ... with its own scratch.
// ---------------------------------------------------------------------------
struct L4fAdjArgs {
    const double2* Lbar;  // [T][N^4]
    const double2* Fs;    // [T][nf3] forward slices
    const double2* Cs;
    const double2* Rs;
    const int* p2f;
    const double2* ops;
    int N, nf, num_ops;
    long nf3;
    long T;
    double2* sbar;        // [T][3][nf3]; zero on entry unless lds_slices
    double2* Abar;        // [T][num_ops][N^2]
    int lds_slices;
    int lds_tb;           // Tbar in LDS (else after U in the scratch)
    double2* scr;         // [gridDim.x][scr_per]
    long scr_per;         // 4 N^2 nf + 3 N^2 + N^4, + N^4 when !lds_tb
    // Gather lists (host-built from the pair map) that give every cotangent
    // entry one owner thread summing its terms in a fixed order:
    // X-form pairs (x, z) with their (a, i, b, form) terms; F and R slice
    // targets with their (n, i, a, b, kind) terms.  Packed: 8 bits per index.
    const int* xp_xz;     // [nxp] x + z * 65536
    const int* xp_start;  // [nxp + 1]
    const int* xp_ent;    // a | i << 8 | b << 16 | form << 24
    int nxp;
    const int* f_tgt;     // [nft] slice index of F targets
    const int* f_start;   // [nft + 1]
    const int* f_ent;     // n | i << 8 | a << 16 | b << 24 (kind in bit 31)
    int nft;
    const int* r_tgt;
    const int* r_start;
    const int* r_ent;
    int nrt;
};

// LDS_SLICES: slice cotangents accumulate in LDS (else in global memory);
// LDS_TB: the N^4 Tbar tile sits in LDS (else in the workgroup's scratch,
// for N above 8 where it exceeds the 160 KiB of a CU).  Both are fixed per
// instantiation so every access has a known address space.
template <bool LDS_SLICES, bool LDS_TB, int BS>
__device__ __forceinline__ void l4f_adjoint_body(double2* sm, const L4fAdjArgs& A) {
    const int N = A.N, N2 = N * N, N4 = N2 * N2, nf = A.nf;
    const long N2nf = static_cast<long>(N2) * nf;
    double2* const scr0 = A.scr + static_cast<long>(blockIdx.x) * A.scr_per;
    double2* Tb = LDS_TB ? sm : scr0 + 4 * N2nf + 3 * N2 + N4;  // N^4
    double2* opsS = LDS_TB ? Tb + N4 : sm;              // num_ops N^2
    double2* Ab = opsS + A.num_ops * N2;                // num_ops N^2
    double2* sb = Ab + A.num_ops * N2;                  // 3 nf3 (LDS_SLICES)
    int* f = reinterpret_cast<int*>(sb + (LDS_SLICES ? 3 * A.nf3 : 0));
    for (int e = threadIdx.x; e < A.num_ops * N2; e += BS) opsS[e] = A.ops[e];
    for (int e = threadIdx.x; e < N2; e += BS) f[e] = A.p2f[e];
    double2* Q = scr0;                                  // Q(a,i,g) at a + i N + g N^2
    double2* S67 = Q + N2nf;
    double2* S67b = S67 + N2nf;
    double2* Qb = S67b + N2nf;
    double2* S123 = Qb + N2nf;                          // N^2
    double2* S123b = S123 + N2;
    double2* Db = S123b + N2;
    double2* U = Db + N2;                               // U(n,i,b,j) at (n + i N) + (b + j N) N^2
    const long nfl = nf;
    auto id3 = [nfl](int x, int y, int z) { return (static_cast<long>(x) * nfl + y) * nfl + z; };
    auto F2 = [f, N](int r, int c) { return f[r + c * N]; };
    for (long t = blockIdx.x; t < A.T; t += gridDim.x) {
        __syncthreads();
        const double2* Lb = A.Lbar + t * N4;
        for (int idx = threadIdx.x; idx < N4; idx += BS) {
            const int ru = idx % N2, cu = idx / N2;
            const int n = ru % N, i = ru / N, m = cu % N, j = cu / N;
            Tb[idx] = cadd(Lb[(n + m * N) + (i + j * N) * N2], cconj(Lb[(m + n * N) + (j + i * N) * N2]));
        }
        for (int e = threadIdx.x; e < A.num_ops * N2; e += BS) Ab[e] = make_double2(0.0, 0.0);
        if constexpr (LDS_SLICES)
            for (long e = threadIdx.x; e < 3 * A.nf3; e += BS) sb[e] = make_double2(0.0, 0.0);
        __syncthreads();
        const double2* Fp = A.Fs + t * A.nf3;
        const double2* Cp = A.Cs + t * A.nf3;
        const double2* Rp = A.Rs + t * A.nf3;
        double2* sF = LDS_SLICES ? sb : A.sbar + t * 3 * A.nf3;
        double2* sC = sF + A.nf3;
        double2* sR = sC + A.nf3;
        for (int iA = 0; iA < A.num_ops; ++iA)
            for (int iB = 0; iB < A.num_ops; ++iB) {
                const double2* Ap = opsS + iA * N2;
                const double2* Bp = opsS + iB * N2;
                double2* AbA = Ab + iA * N2;
                double2* AbB = Ab + iB * N2;
#define AA(r, c) Ap[(r) + (c) * N]
#define BB(r, c) Bp[(r) + (c) * N]
#define TB(n, i, m, j) Tb[((n) + (i) * N) + ((m) + (j) * N) * N2]
                // 1: forward Q and S123
                for (long e = threadIdx.x; e < N2nf; e += BS) {
                    const int a = static_cast<int>(e % N), i = static_cast<int>((e / N) % N);
                    const int g = static_cast<int>(e / N2);
                    double2 y = make_double2(0.0, 0.0);
                    for (int b = 0; b < N; ++b) {
                        const long x1 = id3(F2(b, a), g, F2(a, i)), x2 = id3(F2(i, b), g, F2(a, i));
                        y = cadd(y, cmul(cmul(AA(a, b), BB(b, i)), cadd(Cp[x1], Rp[x1])));
                        y = csub(y, cmul(cmul(BB(a, b), AA(b, i)), cadd(Cp[x2], Rp[x2])));
                    }
                    Q[e] = y;
                }
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int n = e % N, i = e / N;
                    double2 s = make_double2(0.0, 0.0);
                    for (int a = 0; a < N; ++a)
                        for (int b = 0; b < N; ++b) {
                            const double2 Mv = csub(Fp[id3(F2(b, a), F2(b, i), F2(n, b))],
                                                    Rp[id3(F2(b, a), F2(n, a), F2(a, i))]);
                            const double2 Iv = Fp[id3(F2(a, n), F2(b, i), F2(n, b))];
                            const double2 Kv = Rp[id3(F2(i, b), F2(n, a), F2(a, i))];
                            s = csub(s, cmul(cmul(cmul(BB(n, a), AA(a, b)), BB(b, i)), Mv));
                            s = cadd(s, cmul(cmul(cmul(AA(n, a), BB(a, b)), BB(b, i)), Iv));
                            s = csub(s, cmul(cmul(cmul(BB(n, a), BB(a, b)), AA(b, i)), Kv));
                        }
                    S123[e] = s;
                }
                __syncthreads();
                // 2: S67; S123bar, Dbar, S67bar, U from Tbar; A-bar of the A(j,m) S123 term
                for (long e = threadIdx.x; e < N2nf; e += BS) {
                    const int n = static_cast<int>(e % N);
                    const long rest = e - n;
                    double2 s = make_double2(0.0, 0.0);
                    for (int a = 0; a < N; ++a) s = cadd(s, cmul(AA(n, a), Q[rest + a]));
                    S67[e] = s;
                    const int ni = static_cast<int>(e % N2), g = static_cast<int>(e / N2);
                    double2 sb7 = make_double2(0.0, 0.0);
                    for (int j = 0; j < N; ++j)
                        for (int m = 0; m < N; ++m)
                            if (F2(j, m) == g) sb7 = cadd(sb7, cmul(cconj(BB(j, m)), Tb[ni + (m + j * N) * N2]));
                    S67b[e] = sb7;
                }
                for (int e = threadIdx.x; e < N2; e += BS) {
                    double2 s = make_double2(0.0, 0.0), d = s;
                    for (int j = 0; j < N; ++j) {
                        for (int m = 0; m < N; ++m) s = cadd(s, cmul(cconj(AA(j, m)), Tb[e + (m + j * N) * N2]));
                        d = cadd(d, Tb[e + (j + j * N) * N2]);
                    }
                    S123b[e] = s;
                    Db[e] = d;
                }
                for (int e = threadIdx.x; e < N4; e += BS) {
                    const int r = e % N2, c = e / N2, b = c % N, j = c / N;
                    double2 s = make_double2(0.0, 0.0);
                    for (int m = 0; m < N; ++m) s = cadd(s, cmul(cconj(AA(b, m)), Tb[r + (m + j * N) * N2]));
                    U[e] = s;
                }
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int j = e % N, m = e / N;
                    double2 s = make_double2(0.0, 0.0);
                    for (int r = 0; r < N2; ++r) s = cadd(s, cmul(cconj(S123[r]), Tb[r + (m + j * N) * N2]));
                    AbA[j + m * N] = cadd(AbA[j + m * N], s);
                }
                __syncthreads();
                // 3: Qbar; S123bar gets the D term; operator cotangents of the S67,
                // Q and D terms of T and of S67 = A Q and D = -A S123
                for (long e = threadIdx.x; e < N2nf; e += BS) {
                    const int a = static_cast<int>(e % N), i = static_cast<int>((e / N) % N);
                    const int g = static_cast<int>(e / N2);
                    double2 s = make_double2(0.0, 0.0);
                    for (int n = 0; n < N; ++n) s = cadd(s, cmul(cconj(AA(n, a)), S67b[n + i * N + g * N2]));
                    for (int j = 0; j < N; ++j)
                        for (int b = 0; b < N; ++b)
                            if (F2(j, b) == g) s = csub(s, cmul(cconj(BB(j, b)), U[(a + i * N) + (b + j * N) * N2]));
                    Qb[e] = s;
                }
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int a = e % N, i = e / N;
                    double2 s = S123b[e];
                    for (int n = 0; n < N; ++n) s = csub(s, cmul(cconj(AA(n, a)), Db[n + i * N]));
                    S123b[e] = s;
                }
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int j = e % N, m = e / N;
                    const long go = static_cast<long>(F2(j, m)) * N2;
                    double2 s = make_double2(0.0, 0.0);
                    for (int r = 0; r < N2; ++r) s = cadd(s, cmul(cconj(S67[go + r]), Tb[r + (m + j * N) * N2]));
                    AbB[j + m * N] = cadd(AbB[j + m * N], s);
                }
                __syncthreads();
                // v(j,b,m) = sum_r conj(Q(r, f(j,b))) Tbar(r; m,j), in U's space (U is done)
                double2* V = U;
                for (int e = threadIdx.x; e < N2 * N; e += BS) {
                    const int j = e % N, m = e / N2;
                    const long go = static_cast<long>(F2(j, (e / N) % N)) * N2;
                    double2 v = make_double2(0.0, 0.0);
                    for (int r = 0; r < N2; ++r) v = cadd(v, cmul(cconj(Q[go + r]), Tb[r + (m + j * N) * N2]));
                    V[e] = v;
                }
                __syncthreads();
                // the B(j,b) A(b,m) Q term: each cotangent entry sums its v's in j (or m) order
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int b = e % N, m = e / N;
                    double2 s = make_double2(0.0, 0.0);
                    for (int j = 0; j < N; ++j) s = cadd(s, cmul(cconj(BB(j, b)), V[j + b * N + m * N2]));
                    AbA[b + m * N] = csub(AbA[b + m * N], s);
                }
                __syncthreads();
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int j = e % N, b = e / N;
                    double2 s = make_double2(0.0, 0.0);
                    for (int m = 0; m < N; ++m) s = cadd(s, cmul(cconj(AA(b, m)), V[j + b * N + m * N2]));
                    AbB[j + b * N] = csub(AbB[j + b * N], s);
                }
                __syncthreads();
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int n = e % N, a = e / N;
                    double2 s = make_double2(0.0, 0.0);
                    for (int i = 0; i < N; ++i) {
                        s = csub(s, cmul(cconj(S123[a + i * N]), Db[n + i * N]));
                        for (int g = 0; g < nf; ++g)
                            s = cadd(s, cmul(cconj(Q[a + i * N + g * N2]), S67b[n + i * N + g * N2]));
                    }
                    AbA[n + a * N] = cadd(AbA[n + a * N], s);
                }
                __syncthreads();
                // 4: slice and operator cotangents of Q and of S123, each entry owned by
                // one thread that sums its terms in a fixed order (no atomics)
                auto Xv = [&](long x) { return cadd(Cp[x], Rp[x]); };
                // X = C + R slices: pair (x, z) and bucket g gather the Q-bar terms
                for (long e = threadIdx.x; e < static_cast<long>(A.nxp) * nf; e += BS) {
                    const int pp = static_cast<int>(e % A.nxp), g = static_cast<int>(e / A.nxp);
                    double2 acc = make_double2(0.0, 0.0);
                    for (int q = A.xp_start[pp]; q < A.xp_start[pp + 1]; ++q) {
                        const int w = A.xp_ent[q];
                        const int a = w & 255, i = (w >> 8) & 255, b = (w >> 16) & 255;
                        const double2 qb = Qb[a + i * N + g * N2];
                        if (w >> 24)
                            acc = csub(acc, cmul(cconj(cmul(BB(a, b), AA(b, i))), qb));
                        else
                            acc = cadd(acc, cmul(cconj(cmul(AA(a, b), BB(b, i))), qb));
                    }
                    const int xz = A.xp_xz[pp];
                    const long x = id3(xz & 65535, g, xz >> 16);
                    sC[x] = cadd(sC[x], acc);
                    sR[x] = cadd(sR[x], acc);
                }
                __syncthreads();
                // F and R slices of the M, I, K terms of S123
                for (int e = threadIdx.x; e < A.nft; e += BS) {
                    double2 acc = make_double2(0.0, 0.0);
                    for (int q = A.f_start[e]; q < A.f_start[e + 1]; ++q) {
                        const unsigned w = static_cast<unsigned>(A.f_ent[q]);
                        const int n = w & 255, i = (w >> 8) & 255, a = (w >> 16) & 255, b = (w >> 24) & 127;
                        const double2 sv = S123b[n + i * N];
                        if (w >> 31)  // +A(n,a) B(a,b) B(b,i) I
                            acc = cadd(acc, cmul(cconj(cmul(cmul(AA(n, a), BB(a, b)), BB(b, i))), sv));
                        else          // -B(n,a) A(a,b) B(b,i) M, M = F - R
                            acc = csub(acc, cmul(cconj(cmul(cmul(BB(n, a), AA(a, b)), BB(b, i))), sv));
                    }
                    sF[A.f_tgt[e]] = cadd(sF[A.f_tgt[e]], acc);
                }
                for (int e = threadIdx.x; e < A.nrt; e += BS) {
                    double2 acc = make_double2(0.0, 0.0);
                    for (int q = A.r_start[e]; q < A.r_start[e + 1]; ++q) {
                        const unsigned w = static_cast<unsigned>(A.r_ent[q]);
                        const int n = w & 255, i = (w >> 8) & 255, a = (w >> 16) & 255, b = (w >> 24) & 127;
                        const double2 sv = S123b[n + i * N];
                        if (w >> 31)  // -B(n,a) B(a,b) A(b,i) K
                            acc = csub(acc, cmul(cconj(cmul(cmul(BB(n, a), BB(a, b)), AA(b, i))), sv));
                        else          // the -R part of M
                            acc = cadd(acc, cmul(cconj(cmul(cmul(BB(n, a), AA(a, b)), BB(b, i))), sv));
                    }
                    sR[A.r_tgt[e]] = cadd(sR[A.r_tgt[e]], acc);
                }
                __syncthreads();
                auto Mv = [&](int n, int i, int a, int b) {
                    return csub(Fp[id3(F2(b, a), F2(b, i), F2(n, b))], Rp[id3(F2(b, a), F2(n, a), F2(a, i))]);
                };
                auto Iv = [&](int n, int i, int a, int b) { return Fp[id3(F2(a, n), F2(b, i), F2(n, b))]; };
                auto Kv = [&](int n, int i, int a, int b) { return Rp[id3(F2(i, b), F2(n, a), F2(a, i))]; };
                // operator entries at (a,b) of the Q and S123 terms
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int a = e % N, b = e / N;
                    double2 sa = make_double2(0.0, 0.0), sbb = sa;
                    for (int i = 0; i < N; ++i)
                        for (int g = 0; g < nf; ++g) {
                            const double2 qb = Qb[a + i * N + g * N2];
                            sa = cadd(sa, cmul(cconj(cmul(BB(b, i), Xv(id3(F2(b, a), g, F2(a, i))))), qb));
                            sbb = csub(sbb, cmul(cconj(cmul(AA(b, i), Xv(id3(F2(i, b), g, F2(a, i))))), qb));
                        }
                    for (int n = 0; n < N; ++n)
                        for (int i = 0; i < N; ++i) {
                            const double2 sv = S123b[n + i * N];
                            sa = csub(sa, cmul(cconj(cmul(cmul(BB(n, a), BB(b, i)), Mv(n, i, a, b))), sv));
                            sbb = cadd(sbb, cmul(cconj(cmul(cmul(AA(n, a), BB(b, i)), Iv(n, i, a, b))), sv));
                            sbb = csub(sbb, cmul(cconj(cmul(cmul(BB(n, a), AA(b, i)), Kv(n, i, a, b))), sv));
                        }
                    AbA[a + b * N] = cadd(AbA[a + b * N], sa);
                    AbB[a + b * N] = cadd(AbB[a + b * N], sbb);
                }
                __syncthreads();
                // operator entries at (b,i)
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int b = e % N, i = e / N;
                    double2 sa = make_double2(0.0, 0.0), sbb = sa;
                    for (int a = 0; a < N; ++a)
                        for (int g = 0; g < nf; ++g) {
                            const double2 qb = Qb[a + i * N + g * N2];
                            sbb = cadd(sbb, cmul(cconj(cmul(AA(a, b), Xv(id3(F2(b, a), g, F2(a, i))))), qb));
                            sa = csub(sa, cmul(cconj(cmul(BB(a, b), Xv(id3(F2(i, b), g, F2(a, i))))), qb));
                        }
                    for (int n = 0; n < N; ++n)
                        for (int a = 0; a < N; ++a) {
                            const double2 sv = S123b[n + i * N];
                            sbb = csub(sbb, cmul(cconj(cmul(cmul(BB(n, a), AA(a, b)), Mv(n, i, a, b))), sv));
                            sbb = cadd(sbb, cmul(cconj(cmul(cmul(AA(n, a), BB(a, b)), Iv(n, i, a, b))), sv));
                            sa = csub(sa, cmul(cconj(cmul(cmul(BB(n, a), BB(a, b)), Kv(n, i, a, b))), sv));
                        }
                    AbA[b + i * N] = cadd(AbA[b + i * N], sa);
                    AbB[b + i * N] = cadd(AbB[b + i * N], sbb);
                }
                __syncthreads();
                // operator entries at (n,a)
                for (int e = threadIdx.x; e < N2; e += BS) {
                    const int n = e % N, a = e / N;
                    double2 sa = make_double2(0.0, 0.0), sbb = sa;
                    for (int i = 0; i < N; ++i)
                        for (int b = 0; b < N; ++b) {
                            const double2 sv = S123b[n + i * N];
                            sbb = csub(sbb, cmul(cconj(cmul(cmul(AA(a, b), BB(b, i)), Mv(n, i, a, b))), sv));
                            sa = cadd(sa, cmul(cconj(cmul(cmul(BB(a, b), BB(b, i)), Iv(n, i, a, b))), sv));
                            sbb = csub(sbb, cmul(cconj(cmul(cmul(BB(a, b), AA(b, i)), Kv(n, i, a, b))), sv));
                        }
                    AbA[n + a * N] = cadd(AbA[n + a * N], sa);
                    AbB[n + a * N] = cadd(AbB[n + a * N], sbb);
                }
                __syncthreads();
#undef AA
#undef BB
#undef TB
            }
        for (int e = threadIdx.x; e < A.num_ops * N2; e += BS) A.Abar[t * A.num_ops * N2 + e] = Ab[e];
        if constexpr (LDS_SLICES)
            for (long e = threadIdx.x; e < 3 * A.nf3; e += BS) A.sbar[t * 3 * A.nf3 + e] = sb[e];
    }
}

// Every entry is summed by one thread in a fixed order, so the block size
// only changes which thread owns it and the results are identical at any
// size.  Large N (d >= 12: tens of thousands of entries per time, scattered
// slice gathers) runs 1024 threads per time, whose sixteen waves per CU hide
// the gathers' latency (d = 16: 23.2 -> 19.0 ms); small N keeps 256 threads
// (d = 4: 0.43 ms, twice that at 1024).
constexpr int kL4AdjBlockLarge = 1024;
constexpr int kL4AdjLargeN = 12;
template <int BS>
__global__ __launch_bounds__(BS) void k_l4f_adjoint(L4fAdjArgs A) {
    extern __shared__ double2 sm[];
    if (A.lds_tb) {
        if (A.lds_slices)
            l4f_adjoint_body<true, true, BS>(sm, A);
        else
            l4f_adjoint_body<false, true, BS>(sm, A);
    } else {
        if (A.lds_slices)
            l4f_adjoint_body<true, false, BS>(sm, A);
        else
            l4f_adjoint_body<false, false, BS>(sm, A);
    }
}

// ---------------------------------------------------------------------------
// F/C/R adjoint per triple (t < T), the reverse of k_combine and the
// convolutions of phases A-C.  With g1 = Gamma_i, g2 = Gamma_j,
// gm = Gamma_mirror(j), pp = exp(+i Omega t dt), pm = conj(pp), w = g1 pm:
//   F = w PF - convF,  PF = dt cumsum(gm pp),       convF = dt (w * gm)
//   C = w PC - convC,  PC = dt cumsum(conj(g2) pp), convC = dt (w * conj g2)
//   R = g1 P3 - P4,    P3 = dt cumsum(g2 pm),       P4 = dt cumsum(g1 g2 pm)
// (tcl4_kernels.cpp:105-175).  Only t < T enters, so the convolution
// adjoints are direct correlations over [0, T).  Per position p of the
// combine list the scratch holds 11 series of length T:
//   0 w, 1 Fbar, 2 Cbar, 3 wbar, 4 g1bar, 5 g2bar, 6 gmbar, 7..10 P-bars.
// ---------------------------------------------------------------------------
struct FcrAdjArgs {
    const double2* gamma;   // [nf][Nt]
    const double* omegas;
    const int* mirror;
    int nf;
    long Nt;
    double dt;
    long T;
    const int* job_tri;     // combine-list position -> triple
    const unsigned char* job_mask;
    long ncomb;
    const double2* sbar;    // [T][3][nf3]
    long nf3;
    double2* scr;           // [ncomb][11][T]
    double* om;             // [ncomb] Omega cotangent
    const double2* ph_tab;  // phase table of the forward pass (TACO_PHASE_LIBM) or null
    const int* job_row;
    long ph_len;
};

__device__ __forceinline__ double block_sum(double v, double* ws) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0) tot = (ws[0] + ws[1]) + (ws[2] + ws[3]);
    return tot;  // valid in thread 0
}

// Forward prefix scans and reverse suffix scans of one triple.
__global__ __launch_bounds__(kBlock) void k_fcr_adj_scan(FcrAdjArgs P) {
    __shared__ double wsum[8][4];
    __shared__ double wred[4];
    const long p = blockIdx.x;
    const long tri = P.job_tri[p];
    const int mask = P.job_mask[p];
    int i, j, k;
    decode_triple(tri, P.nf, i, j, k);
    const double omega = (P.omegas[i] + P.omegas[j]) + P.omegas[k];
    const double2* prow = P.ph_tab ? P.ph_tab + static_cast<long>(P.job_row[p]) * P.ph_len : nullptr;
    const double2* g1 = P.gamma + static_cast<long>(i) * P.Nt;
    const double2* g2 = P.gamma + static_cast<long>(j) * P.Nt;
    const double2* gm = P.gamma + static_cast<long>(P.mirror[j]) * P.Nt;
    const long T = P.T;
    double2* S = P.scr + p * 11 * T;
    double2 *W = S, *FB = S + T, *CB = S + 2 * T, *WB = S + 3 * T, *G1 = S + 4 * T, *G2 = S + 5 * T,
            *GM = S + 6 * T, *PbF = S + 7 * T, *PbC = S + 8 * T, *Pb3 = S + 9 * T, *Pb4 = S + 10 * T;
    const double dt = P.dt;
    double carry[6] = {0, 0, 0, 0, 0, 0};
    for (long t0 = 0; t0 < T; t0 += kBlock) {
        const long t = t0 + threadIdx.x;
        const bool valid = t < T;
        double2 a1 = make_double2(0, 0), a2 = a1, a3 = a1;
        double2 pp = a1, pm = a1, x1 = a1, x2 = a1;
        if (valid) {
            pp = phase_plus(prow, t, omega, dt);
            pm = cconj(pp);
            x1 = g1[t];
            x2 = g2[t];
            a1 = cmul(gm[t], pp);
            a2 = cmul(cconj(x2), pp);
            a3 = cmul(x2, pm);
        }
        double v[6] = {a1.x, a1.y, a2.x, a2.y, a3.x, a3.y};
        double tot[6];
        block_scan_n<6>(v, wsum, tot);
        for (int q = 0; q < 6; ++q) {
            v[q] = v[q] + carry[q];
            carry[q] += tot[q];
        }
        if (!valid) continue;
        const double2 PF = make_double2(v[0] * dt, v[1] * dt);
        const double2 PC = make_double2(v[2] * dt, v[3] * dt);
        const double2 P3 = make_double2(v[4] * dt, v[5] * dt);
        const double2* sb = P.sbar + t * 3 * P.nf3;
        const double2 z = make_double2(0.0, 0.0);
        const double2 Fb = (mask & 1) ? sb[tri] : z;
        const double2 Cb = (mask & 2) ? sb[P.nf3 + tri] : z;
        const double2 Rb = (mask & 4) ? sb[2 * P.nf3 + tri] : z;
        const double2 w = cmul(x1, pm);
        W[t] = w;
        FB[t] = Fb;
        CB[t] = Cb;
        WB[t] = cadd(cmul(Fb, cconj(PF)), cmul(Cb, cconj(PC)));
        G1[t] = cmul(Rb, cconj(P3));
        PbF[t] = cmul(Fb, cconj(w));
        PbC[t] = cmul(Cb, cconj(w));
        Pb3[t] = cmul(Rb, cconj(x1));
        Pb4[t] = make_double2(-Rb.x, -Rb.y);
    }
    __syncthreads();
    // suffix sums ybar_m = dt sum_{t >= m} Pbar_t, last chunk first
    double sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double om_part = 0.0;
    const long nch = (T + kBlock - 1) / kBlock;
    for (long ch = nch - 1; ch >= 0; --ch) {
        const long t = ch * kBlock + (kBlock - 1 - threadIdx.x);
        const bool valid = t < T;
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (valid) {
            v[0] = PbF[t].x; v[1] = PbF[t].y; v[2] = PbC[t].x; v[3] = PbC[t].y;
            v[4] = Pb3[t].x; v[5] = Pb3[t].y; v[6] = Pb4[t].x; v[7] = Pb4[t].y;
        }
        double tot[8];
        block_scan_n<8>(v, wsum, tot);
        for (int q = 0; q < 8; ++q) {
            v[q] = v[q] + sc[q];
            sc[q] += tot[q];
        }
        if (!valid) continue;
        const double2 yF = make_double2(v[0] * dt, v[1] * dt), yC = make_double2(v[2] * dt, v[3] * dt);
        const double2 y3 = make_double2(v[4] * dt, v[5] * dt), y4 = make_double2(v[6] * dt, v[7] * dt);
        const double tdt = static_cast<double>(t) * dt;
        const double2 pp = phase_plus(prow, t, omega, dt), pm = cconj(pp);
        const double2 x1 = g1[t], x2 = g2[t], xm = gm[t];
        GM[t] = cmul(yF, cconj(pp));
        G2[t] = cadd(cadd(cmul(cconj(yC), pp), cmul(y3, pp)), cmul(y4, cconj(cmul(x1, pm))));
        G1[t] = cadd(G1[t], cmul(y4, cconj(cmul(x2, pm))));
        const double2 ppb = cadd(cmul(yF, cconj(xm)), cmul(yC, x2));
        const double2 pmb = cadd(cmul(y3, cconj(x2)), cmul(y4, cconj(cmul(x1, x2))));
        // d pp / d Omega = i t dt pp, d pm / d Omega = -i t dt pm
        const double2 dpp = make_double2(-tdt * pp.y, tdt * pp.x);
        const double2 dpm = make_double2(tdt * pm.y, -tdt * pm.x);
        om_part += (ppb.x * dpp.x + ppb.y * dpp.y) + (pmb.x * dpm.x + pmb.y * dpm.y);
    }
    const double tot = block_sum(om_part, wred);
    if (threadIdx.x == 0) P.om[p] = tot;
}

// Correlation adjoints of the two convolutions for one block of 256 lags u:
//   wbar_u  += dt sum_{t>=u} (-Fbar_t conj(gm_{t-u}) - Cbar_t g2_{t-u})
//   gmbar_u += dt sum_{t>=u} -Fbar_t conj(w_{t-u})
//   g2bar_u += conj(dt sum_{t>=u} -Cbar_t conj(w_{t-u}))
__global__ __launch_bounds__(kBlock) void k_fcr_adj_corr(FcrAdjArgs P) {
    __shared__ double2 fb[kBlock], cb[kBlock];
    __shared__ double2 wgm[2 * kBlock], wg2[2 * kBlock], ww[2 * kBlock];
    const long p = blockIdx.x;
    const int mask = P.job_mask[p];
    if (!(mask & 3)) return;
    const long tri = P.job_tri[p];
    int i, j, k;
    decode_triple(tri, P.nf, i, j, k);
    const double2* g2 = P.gamma + static_cast<long>(j) * P.Nt;
    const double2* gm = P.gamma + static_cast<long>(P.mirror[j]) * P.Nt;
    const long T = P.T;
    double2* S = P.scr + p * 11 * T;
    const double2 *W = S, *FB = S + T, *CB = S + 2 * T;
    double2 *WB = S + 3 * T, *G2 = S + 5 * T, *GM = S + 6 * T;
    const long ub0 = static_cast<long>(blockIdx.y) * kBlock;
    const long u = ub0 + threadIdx.x;
    double2 xw = make_double2(0, 0), xm = xw, xc = xw;
    for (long t0 = ub0; t0 < T; t0 += kBlock) {
        // window of lags d = t - u for t in [t0, t0 + 256), u in [ub0, ub0 + 256):
        // d in [t0 - ub0 - 255, t0 - ub0 + 255], stored at d - (t0 - ub0 - 255)
        const long dlo = t0 - ub0 - (kBlock - 1);
        for (int e = threadIdx.x; e < 2 * kBlock; e += kBlock) {
            const long d = dlo + e;
            const bool ok = d >= 0 && d < T;
            wgm[e] = ok ? gm[d] : make_double2(0, 0);
            wg2[e] = ok ? g2[d] : make_double2(0, 0);
            ww[e] = ok ? W[d] : make_double2(0, 0);
        }
        {
            const long t = t0 + threadIdx.x;
            fb[threadIdx.x] = t < T ? FB[t] : make_double2(0, 0);
            cb[threadIdx.x] = t < T ? CB[t] : make_double2(0, 0);
        }
        __syncthreads();
        if (u < T) {
            const int qlo = static_cast<int>(max(0L, u - t0));
            const int qhi = static_cast<int>(min(static_cast<long>(kBlock), T - t0));
            for (int q = qlo; q < qhi; ++q) {
                const int e = static_cast<int>(t0 + q - u - dlo);
                const double2 f = fb[q], c = cb[q];
                xw = csub(xw, cadd(cmul(f, cconj(wgm[e])), cmul(c, wg2[e])));
                xm = csub(xm, cmul(f, cconj(ww[e])));
                xc = csub(xc, cmul(c, cconj(ww[e])));
            }
        }
        __syncthreads();
    }
    if (u < T) {
        const double dt = P.dt;
        WB[u] = cadd(WB[u], cscale(xw, dt));
        GM[u] = cadd(GM[u], cscale(xm, dt));
        G2[u] = cadd(G2[u], cconj(cscale(xc, dt)));
    }
}

// g1bar += wbar pp; Omega cotangent of w = g1 pm.
__global__ __launch_bounds__(kBlock) void k_fcr_adj_fin(FcrAdjArgs P) {
    __shared__ double wred[4];
    const long p = blockIdx.x;
    const long tri = P.job_tri[p];
    int i, j, k;
    decode_triple(tri, P.nf, i, j, k);
    const double omega = (P.omegas[i] + P.omegas[j]) + P.omegas[k];
    const double2* prow = P.ph_tab ? P.ph_tab + static_cast<long>(P.job_row[p]) * P.ph_len : nullptr;
    const double2* g1 = P.gamma + static_cast<long>(i) * P.Nt;
    const long T = P.T;
    double2* S = P.scr + p * 11 * T;
    const double2* WB = S + 3 * T;
    double2* G1 = S + 4 * T;
    double om_part = 0.0;
    for (long t = threadIdx.x; t < T; t += kBlock) {
        const double tdt = static_cast<double>(t) * P.dt;
        const double2 pp = phase_plus(prow, t, omega, P.dt), pm = cconj(pp);
        const double2 wb = WB[t];
        G1[t] = cadd(G1[t], cmul(wb, pp));
        const double2 pmb = cmul(wb, cconj(g1[t]));
        const double2 dpm = make_double2(tdt * pm.y, -tdt * pm.x);
        om_part += pmb.x * dpm.x + pmb.y * dpm.y;
    }
    const double tot = block_sum(om_part, wred);
    if (threadIdx.x == 0) P.om[p] += tot;
}

// Gamma cotangent per bucket: gbar[b][t] = sum over the bucket's entries
// (position, role) of series role (4 g1bar, 5 g2bar, 6 gmbar), in list order.
// Per bucket b (blockIdx.y), its list of (position, role) entries in
// gridDim.z contiguous pieces: piece z (blockIdx.z) sums its entries in list
// order into part[(z nf + b) T + t]; k_bucket_reduce_pieces adds the pieces
// in order.  Enough workgroups to hide the gathers' latency (a bucket can
// hold thousands of entries), and a fixed order.
constexpr int kBredPieces = 32;
__global__ __launch_bounds__(kBlock) void k_bucket_reduce(const double2* scr, long T, const int* ent_start,
                                                          const int* ent, double2* part) {
    const int b = blockIdx.y, nf = gridDim.y, z = blockIdx.z, P = gridDim.z;
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= T) return;
    const long q0 = ent_start[b], len = ent_start[b + 1] - q0;
    const long lo = q0 + len * z / P, hi = q0 + len * (z + 1) / P;
    double2 s = make_double2(0.0, 0.0);
    for (long q = lo; q < hi; ++q) {
        const int code = ent[q];
        const long p = code / 3, role = 4 + code % 3;
        s = cadd(s, scr[(p * 11 + role) * T + t]);
    }
    part[(static_cast<long>(z) * nf + b) * T + t] = s;
}

__global__ __launch_bounds__(kBlock) void k_bucket_reduce_pieces(const double2* part, long T, long Nt,
                                                                 double2* gbar) {
    const int b = blockIdx.y, nf = gridDim.y;
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= T) return;
    double2 s = make_double2(0.0, 0.0);
    for (int z = 0; z < kBredPieces; ++z) s = cadd(s, part[(static_cast<long>(z) * nf + b) * T + t]);
    gbar[static_cast<long>(b) * Nt + t] = s;
}

// ---------------------------------------------------------------------------
// TCL2 generator adjoint, the reverse of k_tcl2 (generator.cpp:81-177 with
// diagonal channels).  With P = A masked to bucket b, AdA = P^H P,
// gam = 2 Re Gamma_b(t), S = Im Gamma_b(t) the forward adds per (operator,
// bucket) part gam K (K = dL/dgam, the dissipator direction) and the
// unitary part of H = diag(eps) + hermitize(sum S AdA).  Per time t
// (k_tcl2_adj_t, one workgroup each):
//   Hbar(p,q) = i sum_r Lbar[(r,p),(r,q)] - i sum_r Lbar[(q,r),(p,r)]
//   (row (r1,r2) = r1 N + r2), eps parts Re Hbar(n,n), X = (Hbar + Hbar^H)/2,
//   Gamma_b cotangent (2 sum_parts Re<Lbar,K>, sum_parts Re<X,AdA>).
// Time sums (k_tcl2_adj_w per chunk of times, k_tcl2_adj_wsum over chunks
// in order): W_b = sum_t gam_b Lbar(t), V_b = sum_t S_b X(t) and
// sum_t Re Hbar(n,n); the host forms the operator cotangents from them.
// ---------------------------------------------------------------------------
struct Tcl2AdjArgs {
    const double2* Lbar;   // [T][D^2]
    const double2* gamma;  // [nf][Nt]
    long Nt, T;
    const double2* K;      // [npart][D^2]
    const double2* AdA;    // [npart][N^2]
    const int* part_b;     // [npart] bucket of each part
    int npart, N, nf, order;
    double2* gbar;         // [nf][T]
    double2* X;            // [T][N^2]
    double* epsd;          // [T][N]
    int nchunk;
    long nout;             // nf D^2 + nf N^2 + N
    double2* part;         // [nchunk][nout]
    double2* wsum;         // [nout]: W_b, V_b, eps sums (in .x)
};

// Operator cotangent of every TCL2 part from the time sums W_b (D^2) and
// V_b (N^2) of its bucket: the transposes of the part's K and AdA = P^H P
// terms applied to W_b and V_b,
//   AdAbar(a, c) = V_b(a, c) - 1/2 sum_k W_b[(k N + a) + (k N + c) D]
//                            - 1/2 sum_k W_b[(c N + k) + (a N + k) D],
//   Pbar(x, y) = sum_{r,c} W_b[(r N + x) + (c N + y) D] P(r, c)
//              + sum_{r,c} conj(W_b[(x N + r) + (y N + c) D]) P(r, c)
//              + sum_r AdAbar(r, y) P(x, r) + sum_c conj(AdAbar(y, c)) P(x, c),
// written per part (the host adds Pbar where pair_to_freq = b).  One
// workgroup per part; N^2 <= 1024.
__global__ __launch_bounds__(kBlock) void k_tcl2_ops_parts(const double2* W, const double2* V, const double2* P,
                                                           const int* part_b, int N, double2* Pbar) {
    extern __shared__ double2 shp[];
    const int N2 = N * N, D = N2;
    const long D2 = static_cast<long>(D) * D;
    const long q = blockIdx.x;
    const int b = part_b[q];
    const double2* Wb = W + static_cast<long>(b) * D2;
    const double2* Vb = V + static_cast<long>(b) * N2;
    const double2* Pq = P + q * N2;
    double2* Ps = shp;        // [N2]
    double2* Ab = shp + N2;   // [N2] AdAbar
    for (int e = threadIdx.x; e < N2; e += kBlock) Ps[e] = Pq[e];
    for (int e = threadIdx.x; e < N2; e += kBlock) {
        const int a = e % N, cc = e / N;
        double2 v = Vb[e];
        for (int k = 0; k < N; ++k) v = csub(v, cscale(Wb[(k * N + a) + static_cast<long>(k * N + cc) * D], 0.5));
        for (int k = 0; k < N; ++k) v = csub(v, cscale(Wb[(cc * N + k) + static_cast<long>(a * N + k) * D], 0.5));
        Ab[e] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < N2; e += kBlock) {
        const int x = e % N, y = e / N;
        double2 s = make_double2(0.0, 0.0);
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) s = cadd(s, cmul(Wb[(r * N + x) + static_cast<long>(c * N + y) * D], Ps[r + c * N]));
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r)
                s = cadd(s, cmul(cconj(Wb[(x * N + r) + static_cast<long>(y * N + c) * D]), Ps[r + c * N]));
        for (int r = 0; r < N; ++r) s = cadd(s, cmul(Ab[r + y * N], Ps[x + r * N]));
        for (int c = 0; c < N; ++c) s = cadd(s, cmul(cconj(Ab[y + c * N]), Ps[x + c * N]));
        Pbar[q * N2 + e] = s;
    }
}

// K of every TCL2 part (d L / d gam_b for the part's masked operator P):
// K[(r1 N + r2) + (c1 N + c2) D] = conj(P(r1,c1)) P(r2,c2)
//   - [r1 = c1] AdA(r2,c2)/2 - [r2 = c2] AdA(c1,r1)/2.
// Grid: (element blocks of D^2, parts).
__global__ __launch_bounds__(kBlock) void k_tcl2_part_k(const double2* P, const double2* AdA, int N, double2* K) {
    const int N2 = N * N, D = N2;
    const long D2 = static_cast<long>(D) * D;
    const long e = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (e >= D2) return;
    const long q = blockIdx.y;
    const int row = static_cast<int>(e % D), col = static_cast<int>(e / D);
    const int r1 = row / N, r2 = row % N, c1 = col / N, c2 = col % N;
    const double2* Pq = P + q * N2;
    const double2* Aq = AdA + q * N2;
    double2 v = cmul(cconj(Pq[r1 + c1 * N]), Pq[r2 + c2 * N]);
    if (r1 == c1) v = csub(v, cscale(Aq[r2 + c2 * N], 0.5));
    if (r2 == c2) v = csub(v, cscale(Aq[c1 + r1 * N], 0.5));
    K[q * D2 + e] = v;
}

__global__ __launch_bounds__(kBlock) void k_tcl2_adj_t(Tcl2AdjArgs A) {
    __shared__ double wred[4];
    extern __shared__ double2 sh[];
    const int N = A.N, N2 = N * N, D = N2;
    const long D2 = static_cast<long>(D) * D;
    const long t = blockIdx.x;
    const double2* Lb = A.Lbar + t * D2;
    double2* Hb = sh;
    double2* Xb = sh + N2;
    for (int e = threadIdx.x; e < N2; e += kBlock) {
        const int p = e % N, q = e / N;
        double2 h = make_double2(0.0, 0.0);
        for (int r = 0; r < N; ++r) {
            const double2 a = Lb[(r * N + p) + static_cast<long>(r * N + q) * D];
            const double2 b = Lb[(q * N + r) + static_cast<long>(p * N + r) * D];
            h = cadd(h, make_double2(-a.y, a.x));  // + i a
            h = cadd(h, make_double2(b.y, -b.x));  // - i b
        }
        Hb[e] = h;
    }
    __syncthreads();
    for (int n = threadIdx.x; n < N; n += kBlock) A.epsd[t * N + n] = Hb[n + n * N].x;
    if (A.order == 0) return;
    for (int e = threadIdx.x; e < N2; e += kBlock) {
        const int r = e % N, c = e / N;
        const double2 x = cscale(cadd(Hb[e], cconj(Hb[c + r * N])), 0.5);
        Xb[e] = x;
        A.X[t * N2 + e] = x;
    }
    __syncthreads();
    // Gamma_b cotangent, imaginary part (the real part, 2 sum_parts
    // Re<Lbar,K>, comes from k_tcl2_adj_kdot and k_tcl2_adj_gam)
    for (int b = 0; b < A.nf; ++b) {
        double sp = 0.0;
        bool any = false;
        for (int pi = 0; pi < A.npart; ++pi) {
            if (A.part_b[pi] != b) continue;
            any = true;
            const double2* M = A.AdA + static_cast<long>(pi) * N2;
            for (int e = threadIdx.x; e < N2; e += kBlock) {
                const double2 x = Xb[e], m = M[e];
                sp += x.x * m.x + x.y * m.y;
            }
        }
        double sv = 0.0;
        if (any) sv = block_sum(sp, wred);
        if (threadIdx.x == 0) A.gbar[static_cast<long>(b) * A.T + t] = make_double2(0.0, sv);
    }
}

// G[pi][t] = sum_e Re(conj(K_pi[e]) Lbar_t[e]) over e-chunks (blockIdx.z):
// a workgroup stages 8 parts' and 8 times' rows of 256 entries in LDS and
// each of its 64 (part, time) pairs is summed by 4 threads (entries c = sub
// + 4m of each staged row, in order) that meet in a fixed order; the chunk
// sums land in part[(z npart + pi) T + t].  Each Lbar_t and K_pi row is
// read once per tile of 8 parts / 8 times instead of once per part / time.
constexpr int kKdTile = 8, kKdE = 256;
__global__ __launch_bounds__(kBlock) void k_tcl2_adj_kdot(const double2* K, const double2* Lbar, long D2, int npart,
                                                          long T, long echunk, double* part) {
    __shared__ double2 sk[kKdTile][kKdE], sl[kKdTile][kKdE];
    __shared__ double red[4][64];
    const int p0 = blockIdx.x * kKdTile;
    const long t0 = static_cast<long>(blockIdx.y) * kKdTile;
    const long e0 = static_cast<long>(blockIdx.z) * echunk, e1 = min(D2, e0 + echunk);
    const int pair = threadIdx.x & 63, sub = threadIdx.x >> 6;
    const int pp = pair / kKdTile, tt = pair % kKdTile;
    const double2 zero = make_double2(0.0, 0.0);
    double acc = 0.0;
    for (long eb = e0; eb < e1; eb += kKdE) {
        for (int q = threadIdx.x; q < kKdTile * kKdE; q += kBlock) {
            const int r = q / kKdE, c = q % kKdE;
            const long e = eb + c;
            sk[r][c] = (p0 + r < npart && e < e1) ? K[static_cast<long>(p0 + r) * D2 + e] : zero;
            sl[r][c] = (t0 + r < T && e < e1) ? Lbar[(t0 + r) * D2 + e] : zero;
        }
        __syncthreads();
        for (int c = sub; c < kKdE; c += 4) {
            const double2 k = sk[pp][c], l = sl[tt][c];
            acc += l.x * k.x + l.y * k.y;
        }
        __syncthreads();
    }
    red[sub][pair] = acc;
    __syncthreads();
    if (sub == 0 && p0 + pp < npart && t0 + tt < T)
        part[(static_cast<long>(blockIdx.z) * npart + p0 + pp) * T + t0 + tt] =
            (red[0][pair] + red[1][pair]) + (red[2][pair] + red[3][pair]);
}

// gbar[b][t].x = 2 sum over the parts of bucket b (in order) of their
// chunk sums (in order).
__global__ __launch_bounds__(kBlock) void k_tcl2_adj_gam(const double* part, int npart, int nchunk, long T,
                                                         const int* part_b, double2* gbar) {
    const int b = blockIdx.y;
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= T) return;
    double g = 0.0;
    for (int pi = 0; pi < npart; ++pi) {
        if (part_b[pi] != b) continue;
        for (int z = 0; z < nchunk; ++z) g += part[(static_cast<long>(z) * npart + pi) * T + t];
    }
    gbar[static_cast<long>(b) * T + t].x = 2.0 * g;
}

// Chunk sums over times of W_b, V_b and the eps parts (grid: outputs x chunks).
__global__ __launch_bounds__(kBlock) void k_tcl2_adj_w(Tcl2AdjArgs A) {
    const long o = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (o >= A.nout) return;
    const int N = A.N, N2 = N * N;
    const long D2 = static_cast<long>(N2) * N2;
    const long per = (A.T + A.nchunk - 1) / A.nchunk;
    const long t0 = blockIdx.y * per, t1 = min(A.T, t0 + per);
    double2 s = make_double2(0.0, 0.0);
    const long nW = A.nf * D2, nV = static_cast<long>(A.nf) * N2;
    if (o < nW) {
        if (A.order != 0) {
            const long b = o / D2, e = o % D2;
            for (long t = t0; t < t1; ++t)
                s = cadd(s, cscale(A.Lbar[t * D2 + e], 2.0 * A.gamma[b * A.Nt + t].x));
        }
    } else if (o < nW + nV) {
        if (A.order != 0) {
            const long b = (o - nW) / N2, e = (o - nW) % N2;
            for (long t = t0; t < t1; ++t) s = cadd(s, cscale(A.X[t * N2 + e], A.gamma[b * A.Nt + t].y));
        }
    } else {
        const long n = o - nW - nV;
        for (long t = t0; t < t1; ++t) s.x += A.epsd[t * N + n];
    }
    A.part[static_cast<long>(blockIdx.y) * A.nout + o] = s;
}

__global__ __launch_bounds__(kBlock) void k_tcl2_adj_wsum(Tcl2AdjArgs A) {
    const long o = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (o >= A.nout) return;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < A.nchunk; ++c) s = cadd(s, A.part[static_cast<long>(c) * A.nout + o]);
    A.wsum[o] = s;
}

// ---------------------------------------------------------------------------
// Reverse of the bath chain in front of a tape whose Gamma table the device
// produced: Gamma_b(t_k) = (dt/2) sum_{m=1..k} (phi^{m-1} C_{m-1} + phi^m C_m),
// phi = exp(i w_b dt) (gamma.hpp:166-213), of C = (dw/pi) FFT(S), S the
// sampled KMS spectrum of J(w) (correlation_fft.hpp:154-214).  Every sum runs
// in a fixed order, so the cotangents repeat bit for bit.
// ---------------------------------------------------------------------------

// out = a + b elementwise (either may be null: that part is zero).
__global__ __launch_bounds__(kBlock) void k_gbar_sum(const double2* a, const double2* b, long n, double2* out) {
    const long i = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    double2 s = a ? a[i] : make_double2(0.0, 0.0);
    if (b) s = cadd(s, b[i]);
    out[i] = s;
}

// One workgroup per bucket b.  Gamma_b(k) weighs phi^m C_m by dt/2 for m = 0
// (k >= 1) and m = k, by dt for 0 < m < k, so its cotangent is (dt/2) Q_b(m)
// with Q(0) = sum_{k>=1} Gbar(k) and Q(m) = Gbar(m) + 2 sum_{k>m} Gbar(k).
// Writes Q_b [T] and omega_bar_b = (dt/2) Re sum_m conj(Q(m)) (i m dt) phi^m C_m.
// Each thread owns a contiguous run of m; the runs' suffix offsets come from
// one serial pass over the run totals.
__global__ __launch_bounds__(kBlock) void k_trapz_adj_q(const double2* gbar, long T, const double2* C, double dt,
                                                        const double* omegas, double2* Q, double* wbar) {
    __shared__ double2 part[kBlock];
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x;
    const double2* g = gbar + static_cast<long>(b) * T;
    double2* q = Q + static_cast<long>(b) * T;
    const long seg = (T + kBlock - 1) / kBlock;
    const long m0 = min(T, static_cast<long>(threadIdx.x) * seg), m1 = min(T, m0 + seg);
    double2 s = make_double2(0.0, 0.0);
    for (long m = m1 - 1; m >= m0; --m) s = cadd(s, g[m]);
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 run = make_double2(0.0, 0.0);
        for (int i = kBlock - 1; i >= 0; --i) {
            const double2 p = part[i];
            part[i] = run;
            run = cadd(run, p);
        }
    }
    __syncthreads();
    double2 run = part[threadIdx.x];  // sum_{k >= m1} Gbar(k)
    const double w = omegas[b];
    double acc = 0.0;
    for (long m = m1 - 1; m >= m0; --m) {
        const double2 gm = g[m];
        const double2 qm = m == 0 ? run : make_double2(gm.x + 2.0 * run.x, gm.y + 2.0 * run.y);
        q[m] = qm;
        const double md = static_cast<double>(m) * dt;
        const double2 pc = cmul(cexpi(md * w), C[m]);  // phi^m C_m
        acc += md * (qm.y * pc.x - qm.x * pc.y);        // Re(conj(Q) i pc) m dt
        run = cadd(run, gm);
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) wbar[b] = 0.5 * dt * tot;
}

// x[m] = conj(C_bar(m)) for m < T, C_bar(m) = (dt/2) sum_b conj(phi_b^m) Q_b(m)
// (buckets in order), and 0 for T <= m < n: the input of the forward FFT that
// gives sum_m C_bar(m) exp(+2 pi i k m / n) as its conjugate.
__global__ __launch_bounds__(kBlock) void k_trapz_adj_c(const double2* Q, long T, int nf, double dt,
                                                        const double* omegas, long n, double2* x) {
    const long m = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (m >= n) return;
    double2 s = make_double2(0.0, 0.0);
    if (m < T) {
        const double md = static_cast<double>(m) * dt;
        for (int b = 0; b < nf; ++b) s = cadd(s, cmul(cconj(cexpi(md * omegas[b])), Q[static_cast<long>(b) * T + m]));
    }
    x[m] = make_double2(0.5 * dt * s.x, -(0.5 * dt * s.y));
}

struct BcfAdjArgs {
    const double2* spec;  // X[k] at ((k mod 2^a) << b) + (k >> a)
    int a, b;
    long nfft;
    double scale;         // dw / pi
    double domega, beta, p0, p1;
    int kind;
    double* part;         // [gridDim.x][3]
};

__device__ __forceinline__ double bcf_adj_sbar(const BcfAdjArgs& A, long k) {
    const long idx = ((k & ((1L << A.a) - 1)) << A.b) + (k >> A.a);
    return A.scale * A.spec[idx].x;
}

// J(w) and dJ/dp0, dJ/dp1 for w > 0.
__device__ __forceinline__ void spectral_J_grad(int kind, double p0, double p1, double w, double& J, double& d0,
                                                double& d1) {
    if (kind == 0) {  // p0 w exp(-w / p1)
        const double e = taco_cr::exp_cr(-w / p1, c_exptab);
        J = p0 * w * e;
        d0 = w * e;
        d1 = J * w / (p1 * p1);
    } else {  // 2 p0 p1 w / (w^2 + p1^2)
        const double den = w * w + p1 * p1;
        J = 2.0 * p0 * p1 * w / den;
        d0 = 2.0 * p1 * w / den;
        d1 = 2.0 * p0 * w * (w * w - p1 * p1) / (den * den);
    }
}

// Per k = 1..nfft/2-1 (a thread each): with Sbar = (dw/pi) Re X the spectrum
// cotangent, kms = exp(-beta w), f = (pi/2) / (1 - kms), S_k = f J and
// S_{nfft-k} = f J kms, the terms of dLoss/d(p0, p1, beta); per-workgroup
// sums (block_sum: a fixed shuffle tree) into part.
__global__ __launch_bounds__(kBlock) void k_bcf_adj_terms(BcfAdjArgs A) {
    __shared__ double red[kBlock / 64];
    constexpr double kPiD = 3.14159265358979323846;
    const long half = A.nfft / 2;
    const long k = 1 + static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double g0 = 0.0, g1 = 0.0, gb = 0.0;
    if (k < half) {
        const double sk = bcf_adj_sbar(A, k), sm = bcf_adj_sbar(A, A.nfft - k);
        const double w = A.domega * static_cast<double>(k);
        double J, d0, d1;
        spectral_J_grad(A.kind, A.p0, A.p1, w, J, d0, d1);
        const double kms = taco_cr::exp_cr(-A.beta * w, c_exptab);
        const double f = (kPiD / 2.0) / (1.0 - kms);
        const double sp = f * J;
        const double dsp_db = (kPiD / 2.0) * J * (-w * kms) / ((1.0 - kms) * (1.0 - kms));
        g0 = sk * f * d0 + sm * f * kms * d0;
        g1 = sk * f * d1 + sm * f * kms * d1;
        gb = sk * dsp_db + sm * (dsp_db * kms + sp * (-w * kms));
    }
    g0 = block_sum(g0, red);
    g1 = block_sum(g1, red);
    gb = block_sum(gb, red);
    if (threadIdx.x == 0) {
        A.part[3L * blockIdx.x + 0] = g0;
        A.part[3L * blockIdx.x + 1] = g1;
        A.part[3L * blockIdx.x + 2] = gb;
    }
}

// out[0..2] = the workgroup sums in order; out[3] = Sbar_0, out[4] = Sbar_{nfft/2}.
__global__ void k_bcf_adj_final(BcfAdjArgs A, int nblocks, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < nblocks; ++i) {
        s0 += A.part[3L * i + 0];
        s1 += A.part[3L * i + 1];
        s2 += A.part[3L * i + 2];
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
    out[3] = bcf_adj_sbar(A, 0);
    out[4] = bcf_adj_sbar(A, A.nfft / 2);
}
