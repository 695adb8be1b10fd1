// system_prep.hpp — the system preparation (system.hpp:37-157, tcl4.cpp:40-101)
// as element-level steps shared by the host builder (host_prep.cpp, serial
// loops) and the device builder (engine.hip, a workgroup per system with the
// independent element updates spread over its threads).  Both sides run the
// same floating-point operations on every element in the same order, using
// only +, -, *, / and sqrt (correctly rounded on the host and on the GPU), so
// eigenvalues, eigenvectors, gauge, buckets and A_eig agree bit for bit.
//
// Eigensolver: cyclic complex Jacobi.  The rotation of pair (p, q) with
// a_pq = |a_pq| ph takes tau = (a_qq - a_pp) / (2 |a_pq|) = cot(2 theta) and
// the smaller root t = tan(theta) = sign(tau) / (|tau| + sqrt(1 + tau^2))
// (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.5.2), c = 1 /
// sqrt(1 + t^2), s = t c; G = [[c, s ph], [-s conj(ph), c]] and A <- G^H A G,
// V <- V G.  Sweeps stop when the off-diagonal mass is <= 1e-30 of the total
// (at most 100 sweeps).  Eigenvalues ascend (stable order), and each
// eigenvector is scaled to unit norm with its largest-magnitude entry (the
// first within 1e-14 of the maximum) real positive.
#pragma once

#include <cmath>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define TACO_SYS_HD __host__ __device__
#else
#define TACO_SYS_HD
#endif

namespace taco_sys {

struct Cx {
    double re, im;
};

TACO_SYS_HD inline Cx cx(double r, double i) { return Cx{r, i}; }
TACO_SYS_HD inline Cx mul(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
TACO_SYS_HD inline Cx add(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
TACO_SYS_HD inline Cx conj(Cx a) { return Cx{a.re, -a.im}; }
TACO_SYS_HD inline double norm2(Cx a) { return a.re * a.re + a.im * a.im; }
TACO_SYS_HD inline double magn(Cx a) { return sqrt(norm2(a)); }
TACO_SYS_HD inline Cx scale(Cx a, double s) { return Cx{a.re * s, a.im * s}; }
TACO_SYS_HD inline Cx divr(Cx a, double s) { return Cx{a.re / s, a.im / s}; }

struct Rot {
    Cx gpp, gpq, gqp, gqq;
    bool skip;
};

// Rotation zeroing A(p, q) of the Hermitian A (column-major N x N).
TACO_SYS_HD inline Rot rotation(const Cx* A, int N, int p, int q) {
    Rot g{};
    const Cx apq = A[p + q * N];
    const double mag = magn(apq);
    g.skip = !(mag >= 1e-300);
    if (g.skip) return g;
    const double app = A[p + p * N].re, aqq = A[q + q * N].re;
    const Cx ph = divr(apq, mag);
    const double tau = (aqq - app) / (2.0 * mag);
    double t;
    if (fabs(tau) > 1e150) {
        t = 0.5 / tau;
    } else {
        const double r = sqrt(1.0 + tau * tau);
        t = tau >= 0.0 ? 1.0 / (tau + r) : -1.0 / (r - tau);
    }
    const double c = 1.0 / sqrt(1.0 + t * t);
    const double s = t * c;
    g.gpp = cx(c, 0.0);
    g.gqq = cx(c, 0.0);
    g.gpq = scale(ph, s);
    g.gqp = scale(conj(ph), -s);
    return g;
}

// Column update of row r: (A(r,p), A(r,q)) <- (A(r,p), A(r,q)) G.
TACO_SYS_HD inline void rot_cols(Cx* A, int N, int p, int q, const Rot& g, int r) {
    const Cx arp = A[r + p * N], arq = A[r + q * N];
    A[r + p * N] = add(mul(arp, g.gpp), mul(arq, g.gqp));
    A[r + q * N] = add(mul(arp, g.gpq), mul(arq, g.gqq));
}

// Row update of column c: (A(p,c), A(q,c)) <- G^H (A(p,c), A(q,c)).
TACO_SYS_HD inline void rot_rows(Cx* A, int N, int p, int q, const Rot& g, int c) {
    const Cx apc = A[p + c * N], aqc = A[q + c * N];
    A[p + c * N] = add(mul(conj(g.gpp), apc), mul(conj(g.gqp), aqc));
    A[q + c * N] = add(mul(conj(g.gpq), apc), mul(conj(g.gqq), aqc));
}

// After both updates: the rotated pair is exactly zero, the diagonal real.
TACO_SYS_HD inline void rot_fix(Cx* A, int N, int p, int q) {
    A[p + q * N] = cx(0.0, 0.0);
    A[q + p * N] = cx(0.0, 0.0);
    A[p + p * N] = cx(A[p + p * N].re, 0.0);
    A[q + q * N] = cx(A[q + q * N].re, 0.0);
}

// Sweep test: off-diagonal mass <= 1e-30 of the total (sums in column order).
TACO_SYS_HD inline bool converged(const Cx* A, int N) {
    double off = 0.0, tot = 0.0;
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
            const double v = norm2(A[r + c * N]);
            tot += v;
            if (r != c) off += v;
        }
    return off <= 1e-30 * (tot > 1e-300 ? tot : 1e-300);
}

// Stable ascending order of the diagonal into order[0..N).
TACO_SYS_HD inline void sort_diag(const Cx* A, int N, int* order) {
    for (int i = 0; i < N; ++i) order[i] = i;
    for (int i = 1; i < N; ++i) {
        const int v = order[i];
        const double key = A[v + v * N].re;
        int j = i - 1;
        while (j >= 0 && A[order[j] + order[j] * N].re > key) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = v;
    }
}

// Column src of V, unit norm, largest-magnitude entry real positive, into out.
TACO_SYS_HD inline void gauge_column(const Cx* V, int N, int src, Cx* out) {
    int imax = 0;
    double vmax = -1.0;
    for (int r = 0; r < N; ++r) {
        const double m = magn(V[r + src * N]);
        if (m > vmax + 1e-14) {
            vmax = m;
            imax = r;
        }
    }
    const Cx v0 = V[imax + src * N];
    const double m0 = magn(v0);
    const Cx phase = m0 > 0.0 ? divr(conj(v0), m0) : cx(1.0, 0.0);
    double nrm = 0.0;
    for (int r = 0; r < N; ++r) nrm += norm2(V[r + src * N]);
    nrm = sqrt(nrm);
    for (int r = 0; r < N; ++r) out[r] = divr(mul(V[r + src * N], phase), nrm);
}

// Element (r, c) of U^H A U through tmp = U^H A (two passes, sums over q in order).
TACO_SYS_HD inline Cx uha(const Cx* U, const Cx* A, int N, int r, int c) {
    Cx s = cx(0.0, 0.0);
    for (int q = 0; q < N; ++q) s = add(s, mul(conj(U[q + r * N]), A[q + c * N]));
    return s;
}
TACO_SYS_HD inline Cx tu(const Cx* T, const Cx* U, int N, int r, int c) {
    Cx s = cx(0.0, 0.0);
    for (int q = 0; q < N; ++q) s = add(s, mul(T[r + q * N], U[q + c * N]));
    return s;
}

// build_frequency_buckets (system.hpp:138-157) and the map: pairs (m, n) in
// scan order join the first bucket whose representative (the first omega
// seen) is within tol, buckets sorted by omega afterwards; mirror (closest
// -omega, a bucket within tol of zero is its own mirror) and zero bucket.
// Scratch: rep[N*N] representatives, idx/perm/inv[N*N].  Returns nf.
TACO_SYS_HD inline int buckets(const double* eps, int N, double tol, double* rep, int* idx, int* perm, int* inv,
                               double* omegas, int* mirror, int* p2f, int* zero_index) {
    int nb = 0;
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n) {
            const double w = eps[m] - eps[n];
            int b = nb;
            for (int q = 0; q < nb; ++q)
                if (fabs(rep[q] - w) <= tol) {
                    b = q;
                    break;
                }
            if (b == nb) rep[nb++] = w;
            idx[m + n * N] = b;
        }
    // sort the buckets by representative (distinct keys: any sort gives this order)
    for (int q = 0; q < nb; ++q) perm[q] = q;
    for (int i = 1; i < nb; ++i) {
        const int v = perm[i];
        int j = i - 1;
        while (j >= 0 && rep[perm[j]] > rep[v]) {
            perm[j + 1] = perm[j];
            --j;
        }
        perm[j + 1] = v;
    }
    for (int b = 0; b < nb; ++b) {
        omegas[b] = rep[perm[b]];
        inv[perm[b]] = b;
    }
    for (int e = 0; e < N * N; ++e) p2f[e] = inv[idx[e]];
    const double mtol = tol > 1e-12 ? tol : 1e-12;
    for (int j = 0; j < nb; ++j) {
        const double w = omegas[j];
        if (fabs(w) <= mtol) {
            mirror[j] = j;
            continue;
        }
        double best = INFINITY;
        int bi = -1;
        for (int jp = 0; jp < nb; ++jp) {
            const double dw = fabs(omegas[jp] - (-w));
            if (dw < best) {
                best = dw;
                bi = jp;
            }
        }
        mirror[j] = bi >= 0 ? bi : j;
    }
    int zi = -1;
    for (int b = 0; b < nb && zi < 0; ++b)
        if (fabs(omegas[b]) <= mtol) zi = b;
    if (zi < 0) {
        double best = INFINITY;
        for (int b = 0; b < nb; ++b)
            if (fabs(omegas[b]) < best) {
                best = fabs(omegas[b]);
                zi = b;
            }
    }
    *zero_index = zi;
    return nb;
}

}  // namespace taco_sys
