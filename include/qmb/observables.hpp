// Fourier operators of the C++ facade: A_k = N^-1/2 sum_cells exp(i k.r) A_r for
// a single-site operator A localised on every site of one sublattice.
// Reference: Q/quantum_operators/ArbitraryOperator.cu:248-308 (FourierOperator,
// FullFourierOperator), Q/geometry/base.cpp:7-16 (reciprocal vectors). Their
// expectation values in the ground state come from
// HeisenbergXXZAbelianSymms<T>::ComputeStaticExpValZeroT (ed.hpp).
#pragma once

#include <cmath>
#include <complex>
#include <vector>

#include "lattice.hpp"
#include "operators.hpp"

namespace qmb {

using ManyBodyOperatorC = ManyBodyOperator<std::complex<double>>;

// One operator per sublattice alpha: sum over the cells (i, j) of
// exp(i k.(i a1 + j a2)) A localised on site (i*N2 + j)*M + alpha, times 1/sqrt(N1*N2).
inline std::vector<ManyBodyOperatorC> FourierOperator(const ManyBodyOperatorC &A, const std::vector<double> &k,
                                                      const LatticeGeometryInfo &geom) {
    const int M = static_cast<int>(geom.rx_alpha.size());
    const double a1x = geom.A_mat[0][0], a1y = geom.A_mat[1][0];
    const double a2x = geom.A_mat[0][1], a2y = geom.A_mat[1][1];
    const std::complex<double> norm(1.0 / std::sqrt((double)(geom.N1 * geom.N2)), 0.0);
    std::vector<ManyBodyOperatorC> out(M);
    for (int i = 0; i < geom.N1; ++i)
        for (int j = 0; j < geom.N2; ++j) {
            const double rx = i * a1x + j * a2x, ry = i * a1y + j * a2y;
            const std::complex<double> phase = std::exp(std::complex<double>(0.0, k[0] * rx + k[1] * ry));
            for (int alpha = 0; alpha < M; ++alpha) {
                ManyBodyOperatorC local = A;
                local.LocalizeOperator((i * geom.N2 + j) * M + alpha);
                out[alpha] = out[alpha] + phase * local;
            }
        }
    for (auto &op : out) op = norm * op;
    return out;
}

// FourierOperator at every momentum of the torus, k = i b1/N1 + j b2/N2
// (b1, b2 the reciprocal vectors); neg_sqn takes -k. Indexed [i][j][alpha].
inline std::vector<std::vector<std::vector<ManyBodyOperatorC>>> FullFourierOperator(
    const ManyBodyOperatorC &A, const LatticeGeometryInfo &geom, bool neg_sqn = false) {
    const auto B = ReciprocalBasisVectors(geom);  // columns are b1, b2
    std::vector<std::vector<std::vector<ManyBodyOperatorC>>> out(geom.N1);
    for (int i = 0; i < geom.N1; ++i)
        for (int j = 0; j < geom.N2; ++j) {
            std::vector<double> k = {i * B[0][0] / geom.N1 + j * B[0][1] / geom.N2,
                                     i * B[1][0] / geom.N1 + j * B[1][1] / geom.N2};
            if (neg_sqn) k = {-k[0], -k[1]};
            out[i].push_back(FourierOperator(A, k, geom));
        }
    return out;
}

}  // namespace qmb
