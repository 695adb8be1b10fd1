// adjoint_host.cpp — reverse of the TCL2 generator (host side).
//
// The forward generator is k_tcl2 / the drivers' TCL2 composition
// (generator.cpp:81-177 with diagonal channels, TCL4_spin_boson_example.cpp:
// 293-322): per bucket b and operator al, with P = A_al masked to the pairs
// of bucket b, AdA = P^H P, gam = 2 Re Gamma_b(t), S = Im Gamma_b(t):
//   L[row, col] += gam (conj(P[r1,c1]) P[r2,c2] - 1/2 d(r1,c1) AdA[r2,c2]
//                       - 1/2 d(r2,c2) AdA[c1,r1]),
//   Hls += S AdA,  H = (Hls + Hls^H)/2 + diag(eps),
//   L[row, col] += -i (d(r1,c1) H[r2,c2] - d(r2,c2) H[c1,r1]),
// row = r1 N + r2, col = c1 N + c2, L stored L[row + col D].
// Given Lbar(t) for t < T it returns dLoss/dGamma (rows t < T), and adds the
// cotangents of the operators and of eps.  Complex cotangents follow
// dLoss = Re sum conj(z_bar) dz.
#include <algorithm>
#include <complex>
#include <cstddef>
#include <vector>

#include "engine_internal.hpp"

namespace taco_engine {

using cd = std::complex<double>;

void tcl2_adjoint(const cd* Lbar, std::size_t T, const cd* gamma, std::size_t gld, const double* eps, const cd* ops,
                  int num_ops, int N, const int* p2f, int nf, int order, cd* gamma_bar, std::size_t gbld,
                  cd* ops_bar, double* eps_bar) {
    (void)eps;
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN, D = N2, D2 = D * D;
    const cd I(0.0, 1.0);
    // Per (al, b): P, AdA and the dissipator direction K = dL/dgam.
    struct Part {
        std::vector<cd> P, AdA, K;
        bool any = false;
    };
    std::vector<Part> parts(static_cast<std::size_t>(num_ops) * nf);
    for (int al = 0; al < num_ops; ++al)
        for (int b = 0; b < nf; ++b) {
            Part& q = parts[static_cast<std::size_t>(al) * nf + b];
            q.P.assign(N2, cd(0.0, 0.0));
            const cd* A = ops + static_cast<std::size_t>(al) * N2;
            for (std::size_t e = 0; e < N2; ++e)
                if (p2f[e] == b) {
                    q.P[e] = A[e];
                    q.any = q.any || A[e] != cd(0.0, 0.0);
                }
            if (!q.any) continue;
            q.AdA.assign(N2, cd(0.0, 0.0));
            for (std::size_t r = 0; r < NN; ++r)
                for (std::size_t c = 0; c < NN; ++c) {
                    cd s(0.0, 0.0);
                    for (std::size_t m = 0; m < NN; ++m) s += std::conj(q.P[m + r * NN]) * q.P[m + c * NN];
                    q.AdA[r + c * NN] = s;
                }
            q.K.assign(D2, cd(0.0, 0.0));
            for (std::size_t c1 = 0; c1 < NN; ++c1)
                for (std::size_t c2 = 0; c2 < NN; ++c2)
                    for (std::size_t r1 = 0; r1 < NN; ++r1)
                        for (std::size_t r2 = 0; r2 < NN; ++r2) {
                            const std::size_t row = r1 * NN + r2, col = c1 * NN + c2;
                            cd v = std::conj(q.P[r1 + c1 * NN]) * q.P[r2 + c2 * NN];
                            if (r1 == c1) v -= 0.5 * q.AdA[r2 + c2 * NN];
                            if (r2 == c2) v -= 0.5 * q.AdA[c1 + r1 * NN];
                            q.K[row + col * D] = v;
                        }
        }
    // Accumulated over t: W = sum gam Lbar (per al, b), V = sum S Hlsbar.
    // Times are split into at most 64 fixed blocks, summed in parallel and the
    // block partials added in block order (same values for any thread count).
    const std::size_t nblk = std::min<std::size_t>(T, 64), per = nblk ? (T + nblk - 1) / nblk : 0;
    std::vector<cd> Wp(nblk * parts.size() * D2, cd(0.0, 0.0)), Vp(nblk * parts.size() * N2, cd(0.0, 0.0));
    std::vector<double> epsp(nblk * NN, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (std::ptrdiff_t bk = 0; bk < static_cast<std::ptrdiff_t>(nblk); ++bk) {
        std::vector<cd> Hb(N2), Xb(N2);
        cd* Wb = Wp.data() + static_cast<std::size_t>(bk) * parts.size() * D2;
        cd* Vb = Vp.data() + static_cast<std::size_t>(bk) * parts.size() * N2;
        double* eb = epsp.data() + static_cast<std::size_t>(bk) * NN;
        const std::size_t t_end = std::min(T, (static_cast<std::size_t>(bk) + 1) * per);
        for (std::size_t t = static_cast<std::size_t>(bk) * per; t < t_end; ++t) {
            const cd* Lb = Lbar + t * D2;
            // unitary part: Hbar from Lbar
            for (auto& v : Hb) v = cd(0.0, 0.0);
            for (std::size_t r1 = 0; r1 < NN; ++r1)
                for (std::size_t r2 = 0; r2 < NN; ++r2) {
                    const std::size_t row = r1 * NN + r2;
                    for (std::size_t c2 = 0; c2 < NN; ++c2)  // r1 == c1
                        Hb[r2 + c2 * NN] += I * Lb[row + (r1 * NN + c2) * D];
                    for (std::size_t c1 = 0; c1 < NN; ++c1)  // r2 == c2
                        Hb[c1 + r1 * NN] -= I * Lb[row + (c1 * NN + r2) * D];
                }
            for (std::size_t n = 0; n < NN; ++n) eb[n] += Hb[n + n * NN].real();
            if (order == 0) continue;
            for (std::size_t r = 0; r < NN; ++r)
                for (std::size_t c = 0; c < NN; ++c)
                    Xb[r + c * NN] = 0.5 * (Hb[r + c * NN] + std::conj(Hb[c + r * NN]));
            for (int b = 0; b < nf; ++b) {
                const cd G = gamma[static_cast<std::size_t>(b) * gld + t];
                const double gam = 2.0 * G.real(), S = G.imag();
                double gamb = 0.0, Sb = 0.0;
                for (int al = 0; al < num_ops; ++al) {
                    const std::size_t pi = static_cast<std::size_t>(al) * nf + b;
                    const Part& q = parts[pi];
                    if (!q.any) continue;
                    double g = 0.0;
                    for (std::size_t e = 0; e < D2; ++e) g += (std::conj(Lb[e]) * q.K[e]).real();
                    gamb += g;
                    double sv = 0.0;
                    for (std::size_t e = 0; e < N2; ++e) sv += (std::conj(Xb[e]) * q.AdA[e]).real();
                    Sb += sv;
                    cd* W = Wb + pi * D2;
                    for (std::size_t e = 0; e < D2; ++e) W[e] += gam * Lb[e];
                    cd* V = Vb + pi * N2;
                    for (std::size_t e = 0; e < N2; ++e) V[e] += S * Xb[e];
                }
                // Gamma = gam/2 + i S: dLoss/dRe = 2 gamb, dLoss/dIm = Sb
                gamma_bar[static_cast<std::size_t>(b) * gbld + t] += cd(2.0 * gamb, Sb);
            }
        }
    }
    std::vector<cd> W(parts.size() * D2, cd(0.0, 0.0)), V(parts.size() * N2, cd(0.0, 0.0));
    std::vector<double> epsb(NN, 0.0);
    for (std::size_t bk = 0; bk < nblk; ++bk) {
        for (std::size_t e = 0; e < W.size(); ++e) W[e] += Wp[bk * W.size() + e];
        for (std::size_t e = 0; e < V.size(); ++e) V[e] += Vp[bk * V.size() + e];
        for (std::size_t n = 0; n < NN; ++n) epsb[n] += epsp[bk * NN + n];
    }
    for (std::size_t n = 0; n < NN; ++n) eps_bar[n] += epsb[n];
    if (order == 0) return;
    // Operator cotangents from the accumulated weights.
    std::vector<cd> AdAb(N2), Pb(N2);
    for (int al = 0; al < num_ops; ++al)
        for (int b = 0; b < nf; ++b) {
            const std::size_t pi = static_cast<std::size_t>(al) * nf + b;
            const Part& q = parts[pi];
            if (!q.any) continue;
            const cd* Wq = W.data() + pi * D2;
            const cd* Vq = V.data() + pi * N2;
            for (std::size_t e = 0; e < N2; ++e) AdAb[e] = Vq[e];  // from Hls += S AdA
            for (auto& v : Pb) v = cd(0.0, 0.0);
            for (std::size_t c1 = 0; c1 < NN; ++c1)
                for (std::size_t c2 = 0; c2 < NN; ++c2)
                    for (std::size_t r1 = 0; r1 < NN; ++r1)
                        for (std::size_t r2 = 0; r2 < NN; ++r2) {
                            const std::size_t row = r1 * NN + r2, col = c1 * NN + c2;
                            const cd w = Wq[row + col * D];
                            // jump conj(P[r1,c1]) P[r2,c2]
                            Pb[r2 + c2 * NN] += w * q.P[r1 + c1 * NN];
                            Pb[r1 + c1 * NN] += std::conj(w) * q.P[r2 + c2 * NN];
                            if (r1 == c1) AdAb[r2 + c2 * NN] -= 0.5 * w;
                            if (r2 == c2) AdAb[c1 + r1 * NN] -= 0.5 * w;
                        }
            // AdA[r,c] = sum_m conj(P[m,r]) P[m,c]
            for (std::size_t r = 0; r < NN; ++r)
                for (std::size_t c = 0; c < NN; ++c) {
                    const cd a = AdAb[r + c * NN];
                    for (std::size_t m = 0; m < NN; ++m) {
                        Pb[m + c * NN] += a * q.P[m + r * NN];
                        Pb[m + r * NN] += std::conj(a) * q.P[m + c * NN];
                    }
                }
            cd* Ab = ops_bar + static_cast<std::size_t>(al) * N2;
            for (std::size_t e = 0; e < N2; ++e)
                if (p2f[e] == b) Ab[e] += Pb[e];
        }
}

}  // namespace taco_engine
