// Host check of the Lanczos tridiagonal solver (tridiag_lowest, tridiag_highest
// and tridiag_vector in qmbcuda_amd/csrc/lanczos_host.cpp), which stands in for
// the reference's dense eigensolver call on T
// (Q/exact_diagonalization/lanczos.cu:248-279):
//   - constant tridiagonals, whose spectrum a + 2 b cos(pi j / (m + 1)) and
//     eigenvectors sin(pi i j / (m + 1)) are known in closed form;
//   - random tridiagonals against a dense cyclic Jacobi solve (Golub & Van Loan
//     sec. 8.5) written here: eigenvalues, |last eigenvector component| of the
//     lowest and highest few, ordering, and the eigenvector residual.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "internal.hpp"

using namespace qmb;

static long checked = 0, mismatches = 0;

static void expect(bool ok, const char *what, int m, int j, double got, double want) {
    ++checked;
    if (!ok) {
        ++mismatches;
        if (mismatches <= 20) std::printf("mismatch %s m=%d j=%d got %.17g want %.17g\n", what, m, j, got, want);
    }
}

// Dense symmetric eigensolver by cyclic Jacobi rotations: eigenvalues ascending
// in w, eigenvectors in the columns of V (row-major n x n).
static void jacobi(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (off < 1e-60) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return A[(size_t)x * n + x] < A[(size_t)y * n + y]; });
    std::vector<double> V2((size_t)n * n);
    w.resize(n);
    for (int j = 0; j < n; ++j) {
        w[j] = A[(size_t)ord[j] * n + ord[j]];
        for (int i = 0; i < n; ++i) V2[(size_t)i * n + j] = V[(size_t)i * n + ord[j]];
    }
    V.swap(V2);
}

// ||T z - lam z|| for T = (a, b[1..m-1]).
static double residual(int m, const double *a, const double *b, double lam, const double *z) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) {
        double r = (a[i] - lam) * z[i];
        if (i > 0) r += b[i] * z[i - 1];
        if (i + 1 < m) r += b[i + 1] * z[i + 1];
        s += r * r;
    }
    return std::sqrt(s);
}

static double norm(int m, const double *z) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += z[i] * z[i];
    return std::sqrt(s);
}

static void closed_form(int m, double a0, double b0) {
    const double pi = std::acos(-1.0);
    std::vector<double> a(m, a0), b(m, b0);
    b[0] = 0.0;
    std::vector<double> lam(m), zl(m);
    for (int j = 1; j <= m; ++j) {
        lam[j - 1] = a0 + 2.0 * b0 * std::cos(pi * j / (m + 1));
        zl[j - 1] = std::sqrt(2.0 / (m + 1)) * std::fabs(std::sin(pi * (double)m * j / (m + 1)));
    }
    std::vector<int> ord(m);
    for (int i = 0; i < m; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return lam[x] < lam[y]; });
    const int k = std::min(6, m);
    std::vector<double> w(k), zlast(k), z0(m), wh(k), zh(k), zv(m);
    tridiag_lowest(m, a.data(), b.data(), k, w.data(), zlast.data(), z0.data());
    tridiag_highest(m, a.data(), b.data(), k, wh.data(), zh.data());
    const double tol = 1e-13 * (std::fabs(a0) + 2.0 * std::fabs(b0));
    for (int j = 0; j < k; ++j) {
        const int lo = ord[j], hi = ord[m - 1 - j];
        expect(std::fabs(w[j] - lam[lo]) <= tol, "toeplitz lowest", m, j, w[j], lam[lo]);
        expect(std::fabs(wh[j] - lam[hi]) <= tol, "toeplitz highest", m, j, wh[j], lam[hi]);
        expect(std::fabs(std::fabs(zlast[j]) - zl[lo]) <= 1e-10, "toeplitz lowest zlast", m, j, zlast[j], zl[lo]);
        expect(std::fabs(std::fabs(zh[j]) - zl[hi]) <= 1e-10, "toeplitz highest zlast", m, j, zh[j], zl[hi]);
    }
    // the lowest eigenvector itself: sin(pi i j / (m + 1)) up to sign, first entry >= 0
    const int j0 = ord[0] + 1;
    for (int i = 0; i < m; ++i) {
        double want = std::sqrt(2.0 / (m + 1)) * std::sin(pi * (double)(i + 1) * j0 / (m + 1));
        const double s0 = std::sin(pi * (double)j0 / (m + 1));
        if (s0 < 0) want = -want;
        expect(std::fabs(z0[i] - want) <= 1e-10, "toeplitz z0", m, i, z0[i], want);
    }
    tridiag_vector(m, a.data(), b.data(), lam[ord[m - 1]], zv.data());
    expect(std::fabs(norm(m, zv.data()) - 1.0) <= 1e-12, "toeplitz vector norm", m, 0, norm(m, zv.data()), 1.0);
    expect(residual(m, a.data(), b.data(), lam[ord[m - 1]], zv.data()) <= 1e-12 * (std::fabs(a0) + 2 * std::fabs(b0)),
           "toeplitz vector residual", m, 0, residual(m, a.data(), b.data(), lam[ord[m - 1]], zv.data()), 0.0);
}

static void random_case(std::mt19937_64 &rng, int m) {
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> a(m), b(m, 0.0);
    for (int i = 0; i < m; ++i) a[i] = nd(rng);
    for (int i = 1; i < m; ++i) b[i] = std::fabs(nd(rng)) + 0.05;  // Lanczos betas are positive
    std::vector<double> A((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) {
        A[(size_t)i * m + i] = a[i];
        if (i > 0) A[(size_t)i * m + i - 1] = A[(size_t)(i - 1) * m + i] = b[i];
    }
    std::vector<double> wd, V;
    jacobi(m, A, wd, V);
    double scale = 0.0;
    for (double x : wd) scale = std::max(scale, std::fabs(x));
    const int k = std::min(4, m);
    std::vector<double> w(k), zlast(k), z0(m), wh(k), zh(k), zv(m);
    tridiag_lowest(m, a.data(), b.data(), k, w.data(), zlast.data(), z0.data());
    tridiag_highest(m, a.data(), b.data(), k, wh.data(), zh.data());
    for (int j = 0; j < k; ++j) {
        const int hi = m - 1 - j;
        expect(std::fabs(w[j] - wd[j]) <= 1e-12 * scale, "random lowest", m, j, w[j], wd[j]);
        expect(std::fabs(wh[j] - wd[hi]) <= 1e-12 * scale, "random highest", m, j, wh[j], wd[hi]);
        if (j > 0) expect(w[j] >= w[j - 1] && wh[j] <= wh[j - 1], "ordering", m, j, w[j], w[j - 1]);
        // eigenvectors of an unreduced tridiagonal are unique up to sign (distinct eigenvalues)
        const double vl = std::fabs(V[(size_t)(m - 1) * m + j]), vh = std::fabs(V[(size_t)(m - 1) * m + hi]);
        expect(std::fabs(std::fabs(zlast[j]) - vl) <= 1e-9, "random lowest zlast", m, j, zlast[j], vl);
        expect(std::fabs(std::fabs(zh[j]) - vh) <= 1e-9, "random highest zlast", m, j, zh[j], vh);
        tridiag_vector(m, a.data(), b.data(), wh[j], zv.data());
        const double r = residual(m, a.data(), b.data(), wh[j], zv.data());
        expect(r <= 1e-11 * scale && std::fabs(norm(m, zv.data()) - 1.0) <= 1e-12, "random vector", m, j, r, 0.0);
        double dot = 0.0;
        for (int i = 0; i < m; ++i) dot += zv[i] * V[(size_t)i * m + hi];
        expect(std::fabs(std::fabs(dot) - 1.0) <= 1e-9, "random vector direction", m, j, dot, 1.0);
    }
    double dot = 0.0;
    for (int i = 0; i < m; ++i) dot += z0[i] * V[(size_t)i * m + 0];
    expect(std::fabs(std::fabs(dot) - 1.0) <= 1e-9 && z0[0] >= 0.0, "random z0", m, 0, dot, 1.0);
}

int main() {
    const int sizes[] = {1, 2, 3, 5, 8, 37, 120, 500};
    for (int m : sizes) {
        closed_form(m, 0.3, -0.7);
        closed_form(m, -1.25, 0.5);
    }
    std::mt19937_64 rng(5);
    for (int rep = 0; rep < 40; ++rep)
        for (int m : {1, 2, 3, 4, 7, 16, 33, 60}) random_case(rng, m);
    std::printf("checked %ld values, mismatches %ld\n", checked, mismatches);
    return mismatches ? 1 : 0;
}
