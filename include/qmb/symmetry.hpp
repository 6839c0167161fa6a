// Abelian symmetry groups of the C++ facade: SymmetryClass<T> (character table +
// coordinate maps), the coordinate maps, Tnm_group (translations), Cn_group
// (rotations) and the spin-flip extension of this engine.
// Reference: Q/symmetries/SymmetryClass.h:10-28, SymmetryClass.cu:26-71;
// Groups.h:24-38, Groups.cu:11-113; SymmOperators.cu:6-137;
// Q/utils/misc_funcs.cu:80-90. Coordinates are double (the reference uses an
// Armadillo float matrix).
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <functional>
#include <map>
#include <utility>
#include <stdexcept>
#include <vector>

#include "lattice.hpp"

namespace qmb {

// Transformed coordinates: [0] = x of every site, [1] = y of every site.
using Coords = std::array<std::vector<double>, 2>;
using CoordMap = std::function<Coords(const std::vector<double> &, const std::vector<double> &)>;

inline Coords Iden(const std::vector<double> &R1, const std::vector<double> &R2) { return {R1, R2}; }

// Rotation by 2 pi/n around R0 (SymmOperators.cu:6-18).
inline Coords Cn(const std::vector<double> &R1, const std::vector<double> &R2, int n, std::vector<double> R0) {
    const double th = 2.0 * M_PI / n, c = std::cos(th), s = std::sin(th);
    Coords out{R1, R2};
    for (size_t i = 0; i < R1.size(); ++i) {
        const double x = R1[i] - R0[0], y = R2[i] - R0[1];
        out[0][i] = c * x - s * y + R0[0];
        out[1][i] = s * x + c * y + R0[1];
    }
    return out;
}

// Reflection through the line whose normal makes angle theta (degrees) with x (:21-34).
inline Coords sigma_n_arb(const std::vector<double> &R1, const std::vector<double> &R2, double theta_deg) {
    const double th = theta_deg * M_PI / 180.0, nx = std::cos(th), ny = std::sin(th);
    Coords out{R1, R2};
    for (size_t i = 0; i < R1.size(); ++i) {
        const double p = nx * R1[i] + ny * R2[i];
        out[0][i] = R1[i] - 2 * nx * p;
        out[1][i] = R2[i] - 2 * ny * p;
    }
    return out;
}

// Translation by nx*a1 + ny*a2, folded back into the N1 x N2 cluster through
// the (a1, a2) components (SymmOperators.cu:62-126).
inline Coords Tnm_general(const std::vector<double> &R1, const std::vector<double> &R2, int nx, int ny,
                          const LatticeGeometryInfo &g) {
    const double tol = 0.001;
    const double a1[2] = {g.A_mat[0][0], g.A_mat[1][0]};
    const double a2[2] = {g.A_mat[0][1], g.A_mat[1][1]};
    const double w12 = a1[0] * a2[0] + a1[1] * a2[1];
    const double det = 1.0 - w12 * w12;  // W = [[1, w12], [w12, 1]]
    Coords out{R1, R2};
    for (size_t i = 0; i < R1.size(); ++i) {
        const double x = R1[i] + nx * a1[0] + ny * a2[0];
        const double y = R2[i] + nx * a1[1] + ny * a2[1];
        const double p0 = a1[0] * x + a1[1] * y, p1 = a2[0] * x + a2[1] * y;
        double al0 = (p0 - w12 * p1) / det, al1 = (p1 - w12 * p0) / det;
        al0 = std::fmod(al0, (double)g.N1);
        al1 = std::fmod(al1, (double)g.N2);
        if (std::fabs(al0 - g.N1) < tol) al0 = 0.0;
        if (std::fabs(al1 - g.N2) < tol) al1 = 0.0;
        out[0][i] = al0 * a1[0] + al1 * a2[0];
        out[1][i] = al0 * a1[1] + al1 * a2[1];
    }
    return out;
}

// Coordinates folded back into the N1 x N2 cluster through their (a1, a2)
// components, for point operations that map the cluster onto a periodic image
// of itself (components in [0, N) up to a tolerance of 1e-6).
inline Coords FoldIntoCluster(const std::vector<double> &R1, const std::vector<double> &R2,
                              const LatticeGeometryInfo &g) {
    const double a1[2] = {g.A_mat[0][0], g.A_mat[1][0]};
    const double a2[2] = {g.A_mat[0][1], g.A_mat[1][1]};
    const double det = a1[0] * a2[1] - a1[1] * a2[0];
    Coords out{R1, R2};
    for (size_t i = 0; i < R1.size(); ++i) {
        // solve x = u a1 + v a2 for the cell coordinates (u, v)
        double u = (R1[i] * a2[1] - R2[i] * a2[0]) / det;
        double v = (a1[0] * R2[i] - a1[1] * R1[i]) / det;
        u -= g.N1 * std::floor(u / g.N1 + 1e-6);
        v -= g.N2 * std::floor(v / g.N2 + 1e-6);
        out[0][i] = u * a1[0] + v * a2[0];
        out[1][i] = u * a1[1] + v * a2[1];
    }
    return out;
}

// G2 first, then G1 (SymmOperators.cu:128-137).
inline Coords GetElemProd(const std::vector<double> &R1, const std::vector<double> &R2, const CoordMap &G2,
                          const CoordMap &G1) {
    const Coords a = G2(R1, R2);
    return G1(a[0], a[1]);
}

// exp(2 pi i num/den), exact on quarter turns so that real irreps stay real.
inline std::complex<double> RootOfUnity(long num, long den) {
    num %= den;
    if (num < 0) num += den;
    if ((4 * num) % den == 0) {
        static const std::complex<double> q[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        return q[(4 * num) / den];
    }
    const double ang = 2.0 * M_PI * (double)num / (double)den;
    return {std::cos(ang), std::sin(ang)};
}

template <class T>
class SymmetryClass {
  public:
    SymmetryClass() = default;
    SymmetryClass(std::vector<std::vector<T>> character_table, std::vector<CoordMap> elems)
        : Group_elems(std::move(elems)), char_table(std::move(character_table)) {
        spin_flips.assign(Group_elems.size(), 0);
    }
    virtual ~SymmetryClass() = default;

    std::vector<std::vector<T>> GetCharTable() const { return char_table; }
    std::vector<CoordMap> GetGroupElemVec() const { return Group_elems; }
    std::vector<uint8_t> GetSpinFlips() const { return spin_flips; }

    // group_el_arr[g*LS + dest(s)] = s: the destination table of each element
    // in the reference's convention (SymmetryClass.cu:26-71, misc_funcs.cu:80-90).
    void GetGroupElemArr(const std::vector<double> &R1, const std::vector<double> &R2, uint32_t *group_el_arr) const {
        const int LS = static_cast<int>(R1.size());
        const double tol = 0.001;
        for (size_t g = 0; g < Group_elems.size(); ++g) {
            const Coords t = Group_elems[g](R1, R2);
            for (int s = 0; s < LS; ++s) {
                int q = -1;
                for (int c = 0; c < LS; ++c)
                    if (std::fabs(R1[c] - t[0][s]) < tol && std::fabs(R2[c] - t[1][s]) < tol) {
                        q = c;
                        break;
                    }
                if (q < 0) throw std::runtime_error("GetGroupElemArr: an element maps a site outside the lattice");
                group_el_arr[g * LS + q] = static_cast<uint32_t>(s);
            }
        }
    }

  protected:
    std::vector<CoordMap> Group_elems;
    std::vector<std::vector<T>> char_table;
    std::vector<uint8_t> spin_flips;  // new: 1 where the element also inverts every spin
};

using SymmetryClassC = SymmetryClass<std::complex<double>>;

// Translation group: element e = jx*N2 + jy translates by jx*a1 + jy*a2;
// irrep a = nx*N2 + ny has chi = exp(+2 pi i (nx jx/N1 + ny jy/N2)) (Groups.cu:49-113).
class Tnm_group : public SymmetryClassC {
  public:
    explicit Tnm_group(LatticeGeometryInfo g) : geom_info(std::move(g)) {
        N = geom_info.N1;
        M = geom_info.N2;
        for (int jx = 0; jx < N; ++jx)
            for (int jy = 0; jy < M; ++jy) {
                const LatticeGeometryInfo gi = geom_info;
                Group_elems.push_back([gi, jx, jy](const std::vector<double> &R1, const std::vector<double> &R2) {
                    return Tnm_general(R1, R2, jx, jy, gi);
                });
            }
        for (int nx = 0; nx < N; ++nx)
            for (int ny = 0; ny < M; ++ny) {
                std::vector<std::complex<double>> row;
                for (int jx = 0; jx < N; ++jx)
                    for (int jy = 0; jy < M; ++jy)
                        row.push_back(RootOfUnity(((long)(nx * jx) % N) * M + ((long)(ny * jy) % M) * N, (long)N * M));
                char_table.push_back(row);
            }
        spin_flips.assign(Group_elems.size(), 0);
    }
    int get_n() const { return N; }
    int get_m() const { return M; }

  protected:
    int N = 0, M = 0;
    LatticeGeometryInfo geom_info;
};

// Rotation group C_n about R0: element i = rotation by 2 pi i/n,
// chi_a(i) = exp(2 pi i a i/n) (Groups.cu:11-43).
class Cn_group : public SymmetryClassC {
  public:
    Cn_group(int n_in, std::vector<double> R0_in) : n(n_in), R0(std::move(R0_in)) {
        for (int i = 0; i < n; ++i) {
            const int steps = i;
            const auto r0 = R0;
            const int nn = n;
            Group_elems.push_back([steps, r0, nn](const std::vector<double> &R1, const std::vector<double> &R2) {
                Coords c{R1, R2};
                for (int k = 0; k < steps; ++k) c = Cn(c[0], c[1], nn, r0);
                return c;
            });
        }
        for (int a = 0; a < n; ++a) {
            std::vector<std::complex<double>> row;
            for (int i = 0; i < n; ++i) row.push_back(RootOfUnity((long)a * i, n));
            char_table.push_back(row);
        }
        spin_flips.assign(Group_elems.size(), 0);
    }
    int get_n() const { return n; }

  protected:
    int n;
    std::vector<double> R0;
};

// Product set {a o b} (b first) of two groups acting on one lattice, e.g. the
// translations with a reflection or a rotation about a site (new in this
// engine; the reference uses point groups only without translations). With
// `fold`, coordinates are folded back into that torus cluster after each
// factor. Element ia*|B| + ib, spin flips add; characters chi_i(a) chi_j(b) for
// the pairs (i, j) that are one-dimensional representations of the product on
// the sites R1, R2 (a momentum the point operation does not fix gives none),
// kept in (i, j) order; GetIrrepPairs() lists them.
template <class T>
class ProductGroup : public SymmetryClass<T> {
  public:
    ProductGroup(const SymmetryClass<T> &A, const SymmetryClass<T> &B, const std::vector<double> &R1,
                 const std::vector<double> &R2, const LatticeGeometryInfo *fold = nullptr) {
        const auto ea = A.GetGroupElemVec(), eb = B.GetGroupElemVec();
        auto fa = A.GetSpinFlips(), fb = B.GetSpinFlips();
        if (fa.empty()) fa.assign(ea.size(), 0);
        if (fb.empty()) fb.assign(eb.size(), 0);
        const size_t GA = ea.size(), GB = eb.size(), G = GA * GB;
        const bool has_fold = fold != nullptr;
        const LatticeGeometryInfo geom = has_fold ? *fold : LatticeGeometryInfo();
        for (size_t ia = 0; ia < GA; ++ia)
            for (size_t ib = 0; ib < GB; ++ib) {
                const CoordMap a = ea[ia], b = eb[ib];
                this->Group_elems.push_back([a, b, has_fold, geom](const std::vector<double> &X,
                                                                   const std::vector<double> &Y) {
                    Coords c = b(X, Y);
                    if (has_fold) c = FoldIntoCluster(c[0], c[1], geom);
                    c = a(c[0], c[1]);
                    if (has_fold) c = FoldIntoCluster(c[0], c[1], geom);
                    return c;
                });
                this->spin_flips.push_back(static_cast<uint8_t>(fa[ia] ^ fb[ib]));
            }
        // composition table on these sites: source tables ge[g][dest] = src
        const size_t LS = R1.size();
        std::vector<uint32_t> ge(G * LS);
        this->GetGroupElemArr(R1, R2, ge.data());
        std::map<std::pair<std::vector<uint32_t>, int>, size_t> index;
        for (size_t e = 0; e < G; ++e)
            index.emplace(std::make_pair(std::vector<uint32_t>(ge.begin() + e * LS, ge.begin() + (e + 1) * LS),
                                         (int)this->spin_flips[e]),
                          e);
        std::vector<size_t> comp(G * G);
        std::vector<uint32_t> pr(LS);
        for (size_t x = 0; x < G; ++x)
            for (size_t y = 0; y < G; ++y) {
                for (size_t d = 0; d < LS; ++d) pr[d] = ge[y * LS + ge[x * LS + d]];
                auto it = index.find({pr, (int)(this->spin_flips[x] ^ this->spin_flips[y])});
                if (it == index.end()) throw std::runtime_error("ProductGroup: the product is not closed");
                comp[x * G + y] = it->second;
            }
        const auto ca = A.GetCharTable(), cb = B.GetCharTable();
        for (size_t i = 0; i < ca.size(); ++i)
            for (size_t j = 0; j < cb.size(); ++j) {
                std::vector<T> row(G);
                for (size_t e = 0; e < G; ++e) row[e] = ca[i][e / GB] * cb[j][e % GB];
                bool hom = true;
                for (size_t x = 0; x < G && hom; ++x)
                    for (size_t y = 0; y < G && hom; ++y)
                        hom = std::abs(row[comp[x * G + y]] - row[x] * row[y]) < 1e-9;
                if (!hom) continue;
                this->char_table.push_back(row);
                pairs.push_back({(int)i, (int)j});
            }
    }
    std::vector<std::pair<int, int>> GetIrrepPairs() const { return pairs; }

  private:
    std::vector<std::pair<int, int>> pairs;
};

// Z2 spin inversion times a group (new in this engine; the reference has no
// spin flip). Element f*G + e = element e followed by flipping every spin if
// f = 1; irrep p*NIr + a has character chi_a(e) * (-1)^(p*f).
template <class T>
class SpinFlipGroup : public SymmetryClass<T> {
  public:
    explicit SpinFlipGroup(const SymmetryClass<T> &base) {
        const auto el = base.GetGroupElemVec();
        const auto ch = base.GetCharTable();
        const auto bf = base.GetSpinFlips();
        const size_t G = el.size();
        for (int f = 0; f < 2; ++f)
            for (size_t e = 0; e < G; ++e) {
                this->Group_elems.push_back(el[e]);
                this->spin_flips.push_back(static_cast<uint8_t>((bf.empty() ? 0 : bf[e]) ^ f));
            }
        for (int p = 0; p < 2; ++p)
            for (const auto &row : ch) {
                std::vector<T> r;
                for (int f = 0; f < 2; ++f)
                    for (size_t e = 0; e < G; ++e) r.push_back((p && f) ? -row[e] : row[e]);
                this->char_table.push_back(r);
            }
    }
};

}  // namespace qmb
