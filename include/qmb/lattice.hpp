// Geometry and lattice models of the C++ facade: LatticeGeometryInfo,
// Lattice, SpinHamiltonian<T>, HeisenbergInfo<T>, Heisenberg<T> and the
// extended-XXX factories.
// Reference: Q/geometry/base.h:5-19, base.cpp:7-87; Q/lattice_models/Lattice.cpp:6-89,
// ArbitrarySpinLattice.cu:7-61, Heisenberg.h:10-45, Heisenberg.cu:9-214.
#pragma once

#include <algorithm>
#include <cmath>
#include <utility>
#include <stdexcept>
#include <vector>

#include "operators.hpp"

namespace qmb {

struct LatticeGeometryInfo {
    std::vector<double> rx_alpha;             // x-coordinates of the sublattice sites in a cell
    std::vector<double> ry_alpha;             // y-coordinates
    std::vector<std::vector<double>> A_mat;   // columns are the primitive vectors a1, a2
    int N1 = 0;                               // cells along a1
    int N2 = 0;                               // cells along a2
};

inline LatticeGeometryInfo create_square_lattice_info(int N1, int N2) {
    LatticeGeometryInfo g;
    g.A_mat = {{1.0, 0.0}, {0.0, 1.0}};
    g.N1 = N1;
    g.N2 = N2;
    g.rx_alpha = {0.0};
    g.ry_alpha = {0.0};
    return g;
}

inline LatticeGeometryInfo create_triangular_lattice_info(int N1, int N2) {
    LatticeGeometryInfo g;
    const double sq3 = std::sqrt(3.0) / 2.0;
    g.A_mat = {{1.0, 0.5}, {0.0, sq3}};
    g.N1 = N1;
    g.N2 = N2;
    g.rx_alpha = {0.0};
    g.ry_alpha = {0.0};
    return g;
}

inline std::vector<std::vector<double>> ReciprocalBasisVectors(const LatticeGeometryInfo &g) {
    const auto &A = g.A_mat;
    const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const double c = 2.0 * M_PI / det;
    return {{c * A[1][1], -c * A[1][0]}, {-c * A[0][1], c * A[0][0]}};
}

class Lattice {
  public:
    Lattice() = default;
    Lattice(std::vector<double> R1_in, std::vector<double> R2_in) : R1(std::move(R1_in)), R2(std::move(R2_in)) {
        LS = static_cast<int>(R1.size());
    }
    virtual ~Lattice() = default;
    std::vector<double> GetR1() const { return R1; }
    std::vector<double> GetR2() const { return R2; }
    int GetLS() const { return LS; }
    std::vector<double> GetCentralPoint() const {
        double x = 0, y = 0;
        for (int i = 0; i < LS; ++i) {
            x += R1[i];
            y += R2[i];
        }
        return {x / LS, y / LS};
    }

  protected:
    std::vector<double> R1, R2;
    int LS = 0;
};

template <class T>
class SpinHamiltonian : public Lattice {
  public:
    SpinHamiltonian() = default;
    SpinHamiltonian(LatticeGeometryInfo g, ManyBodyOperator<T> H_in) : H(std::move(H_in)), geom_info(std::move(g)) {
        CreateGeometry();
    }
    // Site (i, j, alpha) -> index (i*N2 + j)*M + alpha at i*a1 + j*a2 + r_alpha
    // (ArbitrarySpinLattice.cu:19-36).
    void CreateGeometry() {
        R1.clear();
        R2.clear();
        const double a1x = geom_info.A_mat[0][0], a1y = geom_info.A_mat[1][0];
        const double a2x = geom_info.A_mat[0][1], a2y = geom_info.A_mat[1][1];
        for (int i = 0; i < geom_info.N1; ++i)
            for (int j = 0; j < geom_info.N2; ++j)
                for (size_t a = 0; a < geom_info.rx_alpha.size(); ++a) {
                    R1.push_back(i * a1x + j * a2x + geom_info.rx_alpha[a]);
                    R2.push_back(i * a1y + j * a2y + geom_info.ry_alpha[a]);
                }
        LS = static_cast<int>(R1.size());
    }
    ManyBodyOperator<T> GetH() const { return H; }
    void SetH(ManyBodyOperator<T> H_in) { H = std::move(H_in); }
    LatticeGeometryInfo GetGeomInfo() const { return geom_info; }

  protected:
    ManyBodyOperator<T> H;
    LatticeGeometryInfo geom_info;
};

// Rows of Jxy_terms / Jz_terms are (di, dj, a, b): target cell offset, target
// sublattice a, source sublattice b (Heisenberg.h:10-30).
template <class T>
struct HeisenbergInfo {
    std::vector<std::vector<int>> Jxy_terms;
    std::vector<T> Jxy_couplings;
    std::vector<std::vector<int>> Jz_terms;
    std::vector<double> Jz_couplings;
    bool intra_c_bool = true;
    bool inter_c_bool = true;
};

template <class T>
class Heisenberg : public SpinHamiltonian<T> {
  public:
    Heisenberg() = default;
    Heisenberg(LatticeGeometryInfo g, HeisenbergInfo<T> info) {
        this->geom_info = std::move(g);
        this->CreateGeometry();
        heisenberg_info = std::move(info);
        CreateHamiltonian();
    }
    HeisenbergInfo<T> GetInfo() const { return heisenberg_info; }

    // For every cell (row-major) the Jxy rows (J * Sp_i * Sm_j) then the Jz rows
    // (Jz * Sz_i * Sz_j); boundary filters as in Heisenberg.cu:36-58, including the
    // reference's inter-cluster condition, which is true for every term.
    void CreateHamiltonian() {
        const auto &g = this->geom_info;
        const int M = static_cast<int>(g.rx_alpha.size()), N1 = g.N1, N2 = g.N2;
        this->H = ManyBodyOperator<T>();
        const bool intra = heisenberg_info.intra_c_bool, inter = heisenberg_info.inter_c_bool;
        if (!intra && !inter)
            throw std::invalid_argument(
                "Heisenberg::CreateHamiltonian: intra_c_bool and inter_c_bool are both false, so no bond is kept");
        const std::vector<std::vector<int>> *rows[2] = {&heisenberg_info.Jxy_terms, &heisenberg_info.Jz_terms};
        for (int uc = 0; uc < N1 * N2; ++uc) {
            const int j1 = uc / N2, j2 = uc % N2;
            for (int kind = 0; kind < 2; ++kind) {
                for (size_t ji = 0; ji < rows[kind]->size(); ++ji) {
                    const auto &r = (*rows[kind])[ji];
                    const int j_site = uc * M + r[3];
                    int i1 = j1 + r[0], i2 = j2 + r[1];
                    if (!intra && ((i1 < 0 || i1 >= N1) || (i2 < 0 || i2 >= N2))) continue;
                    if (!inter && ((i1 >= 0 || i1 < N1) || (i2 >= 0 || i2 < N2))) continue;
                    i1 = ((i1 % N1) + N1) % N1;
                    i2 = ((i2 % N2) + N2) % N2;
                    const int i_site = r[2] + M * (N2 * i1 + i2);
                    if (kind == 0)
                        this->H = this->H + (Sp<T>(i_site) * Sm<T>(j_site)) * heisenberg_info.Jxy_couplings[ji];
                    else
                        this->H = this->H + (Sz<T>(i_site) * Sz<T>(j_site)) * T(heisenberg_info.Jz_couplings[ji]);
                }
            }
        }
    }

  protected:
    HeisenbergInfo<T> heisenberg_info;
};

// Every (ni, nj, a, b) with ni in [-N1, N1), nj in [-N2, N2), stable-sorted by
// bond length; shells of equal length (1e-3) get J_vec[0], J_vec[1], ...; both
// Jxy and Jz (Heisenberg.cu:88-176). Distances in float32 like the reference
// so shell boundaries agree.
template <class T>
Heisenberg<T> CreateExtendedHeisenbergXXX(LatticeGeometryInfo g, std::vector<double> J_vec, bool intra = true,
                                          bool inter = true) {
    const int M = static_cast<int>(g.rx_alpha.size());
    std::vector<std::vector<int>> rows;
    std::vector<std::pair<int, float>> dist;
    int idx = 0;
    const auto &A = g.A_mat;
    for (int ni = -g.N1; ni < g.N1; ++ni)
        for (int nj = -g.N2; nj < g.N2; ++nj)
            for (int ai = 0; ai < M; ++ai)
                for (int bi = 0; bi < M; ++bi) {
                    rows.push_back({ni, nj, ai, bi});
                    const float dx = (float)ni * (float)A[0][0] + (float)nj * (float)A[0][1] + (float)g.rx_alpha[ai] -
                                     (float)g.rx_alpha[bi];
                    const float dy = (float)ni * (float)A[1][0] + (float)nj * (float)A[1][1] + (float)g.ry_alpha[ai] -
                                     (float)g.ry_alpha[bi];
                    // pow(float, int) promotes to double in the reference; sqrtf of the float sum
                    const double d2 = (double)dx * (double)dx + (double)dy * (double)dy;
                    dist.emplace_back(idx++, std::sqrt((float)d2));
                }
    std::stable_sort(dist.begin(), dist.end(), [](const auto &a, const auto &b) { return a.second < b.second; });
    HeisenbergInfo<T> info;
    double J_cur = J_vec[0];
    int J_counter = -1;
    float d_cur = dist[0].second;
    const float thold = 0.001f;
    for (const auto &[k, d] : dist) {
        if (d < thold) continue;
        if (d > d_cur + thold) {
            ++J_counter;
            if (J_counter >= static_cast<int>(J_vec.size())) break;
            J_cur = J_vec[J_counter];
            d_cur = d;
        }
        info.Jxy_terms.push_back(rows[k]);
        info.Jz_terms.push_back(rows[k]);
        info.Jxy_couplings.push_back(T(J_cur));
        info.Jz_couplings.push_back(J_cur);
    }
    info.intra_c_bool = intra;
    info.inter_c_bool = inter;
    return Heisenberg<T>(g, info);
}

template <class T>
Heisenberg<T> CreateHeisenbergXXXSquare(int N1, int N2, std::vector<double> J_vec, bool intra = true,
                                        bool inter = true) {
    return CreateExtendedHeisenbergXXX<T>(create_square_lattice_info(N1, N2), J_vec, intra, inter);
}

template <class T>
Heisenberg<T> CreateHeisenbergXXXTriangular(int N1, int N2, std::vector<double> J_vec, bool intra = true,
                                            bool inter = true) {
    return CreateExtendedHeisenbergXXX<T>(create_triangular_lattice_info(N1, N2), J_vec, intra, inter);
}

}  // namespace qmb
