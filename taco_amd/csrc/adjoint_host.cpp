// adjoint_host.cpp — host side of the TCL2 generator reverse.
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
// The per-time work of the reverse runs on the device (k_tcl2_adj_*); here
// are the per-(operator, bucket) parts it needs and the operator cotangents
// formed from its time sums W_b = sum_t gam_b Lbar(t), V_b = sum_t S_b X(t).
// Complex cotangents follow dLoss = Re sum conj(z_bar) dz.
#include <complex>
#include <cstddef>
#include <vector>

#include "engine_internal.hpp"

namespace taco_engine {

using cd = std::complex<double>;

Tcl2Parts tcl2_parts(const cd* ops, int num_ops, int N, const int* p2f, int nf) {
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN;
    Tcl2Parts out;
    for (int al = 0; al < num_ops; ++al)
        for (int b = 0; b < nf; ++b) {
            std::vector<cd> P(N2, cd(0.0, 0.0));
            const cd* A = ops + static_cast<std::size_t>(al) * N2;
            bool any = false;
            for (std::size_t e = 0; e < N2; ++e)
                if (p2f[e] == b) {
                    P[e] = A[e];
                    any = any || A[e] != cd(0.0, 0.0);
                }
            if (!any) continue;
            std::vector<cd> AdA(N2, cd(0.0, 0.0));
            for (std::size_t r = 0; r < NN; ++r)
                for (std::size_t c = 0; c < NN; ++c) {
                    cd s(0.0, 0.0);
                    for (std::size_t m = 0; m < NN; ++m) s += std::conj(P[m + r * NN]) * P[m + c * NN];
                    AdA[r + c * NN] = s;
                }
            out.P.insert(out.P.end(), P.begin(), P.end());
            out.AdA.insert(out.AdA.end(), AdA.begin(), AdA.end());
            out.al.push_back(al);
            out.b.push_back(b);
        }
    return out;
}

void tcl2_ops_adjoint(const Tcl2Parts& parts, const cd* W, const cd* V, int N, const int* p2f, cd* ops_bar) {
    const std::size_t NN = static_cast<std::size_t>(N), N2 = NN * NN, D = N2, D2 = D * D;
    std::vector<cd> AdAb(N2), Pb(N2);
    for (std::size_t q = 0; q < parts.al.size(); ++q) {
        const int b = parts.b[q];
        const cd* P = parts.P.data() + q * N2;
        const cd* Wq = W + static_cast<std::size_t>(b) * D2;
        const cd* Vq = V + static_cast<std::size_t>(b) * N2;
        for (std::size_t e = 0; e < N2; ++e) AdAb[e] = Vq[e];  // from Hls += S AdA
        for (auto& v : Pb) v = cd(0.0, 0.0);
        for (std::size_t c1 = 0; c1 < NN; ++c1)
            for (std::size_t c2 = 0; c2 < NN; ++c2)
                for (std::size_t r1 = 0; r1 < NN; ++r1)
                    for (std::size_t r2 = 0; r2 < NN; ++r2) {
                        const std::size_t row = r1 * NN + r2, col = c1 * NN + c2;
                        const cd w = Wq[row + col * D];
                        // jump conj(P[r1,c1]) P[r2,c2]
                        Pb[r2 + c2 * NN] += w * P[r1 + c1 * NN];
                        Pb[r1 + c1 * NN] += std::conj(w) * P[r2 + c2 * NN];
                        if (r1 == c1) AdAb[r2 + c2 * NN] -= 0.5 * w;
                        if (r2 == c2) AdAb[c1 + r1 * NN] -= 0.5 * w;
                    }
        // AdA[r,c] = sum_m conj(P[m,r]) P[m,c]
        for (std::size_t r = 0; r < NN; ++r)
            for (std::size_t c = 0; c < NN; ++c) {
                const cd a = AdAb[r + c * NN];
                for (std::size_t m = 0; m < NN; ++m) {
                    Pb[m + c * NN] += a * P[m + r * NN];
                    Pb[m + r * NN] += std::conj(a) * P[m + c * NN];
                }
            }
        cd* Ab = ops_bar + static_cast<std::size_t>(parts.al[q]) * N2;
        for (std::size_t e = 0; e < N2; ++e)
            if (p2f[e] == b) Ab[e] += Pb[e];
    }
}

}  // namespace taco_engine
