// Host FP64 tridiagonal eigensolver for the Lanczos matrix T.
//
// Stands in for the reference's FP32 dense cusolverDnXsyevd call on the
// NIter x NIter tridiagonal (Q/cuda_wrappers/CudaWrappers.cu:174-266,
// Q/exact_diagonalization/lanczos.cu:248-279). Only the lowest few eigenpairs
// are needed, so the dense matrix is never formed:
//   eigenvalues  : Sturm-sequence bisection (Wilkinson, The Algebraic
//                  Eigenvalue Problem, ch. 5; Golub & Van Loan sec. 8.4.1);
//   eigenvectors : inverse iteration with a pivoted tridiagonal LU.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace qmb {

namespace {

// Number of eigenvalues of T strictly below x (LDL^T inertia).
int sturm_count(int m, const double *a, const double *b, double x, double pivmin) {
    int cnt = 0;
    double q = a[0] - x;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++cnt;
    for (int i = 1; i < m; ++i) {
        q = a[i] - x - b[i] * b[i] / q;
        if (std::fabs(q) < pivmin) q = -pivmin;
        if (q < 0) ++cnt;
    }
    return cnt;
}

double bisect(int m, const double *a, const double *b, int j, double lo, double hi, double pivmin) {
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (hi - lo <= 2.0 * DBL_EPSILON * std::max(std::fabs(lo), std::fabs(hi)) + pivmin) break;
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(m, a, b, mid, pivmin) > j)
            hi = mid;
        else
            lo = mid;
    }
    return 0.5 * (lo + hi);
}

// Solves (T - lam I) x = rhs in place with a pivoted LU (like LAPACK's gttrf/gtts2).
void shifted_solve(int m, const double *a, const double *b, double lam, double tiny, std::vector<double> &x) {
    std::vector<double> d(m), du(m, 0.0), dl(m, 0.0), du2(m, 0.0);
    std::vector<int> piv(m, 0);
    for (int i = 0; i < m; ++i) d[i] = a[i] - lam;
    for (int i = 0; i + 1 < m; ++i) {
        du[i] = b[i + 1];
        dl[i] = b[i + 1];
    }
    for (int i = 0; i + 1 < m; ++i) {
        if (std::fabs(d[i]) >= std::fabs(dl[i])) {
            if (d[i] == 0.0) d[i] = tiny;
            const double f = dl[i] / d[i];
            dl[i] = f;
            d[i + 1] -= f * du[i];
            du2[i] = 0.0;
            piv[i] = 0;
        } else {
            const double f = d[i] / dl[i];
            d[i] = dl[i];
            dl[i] = f;
            const double t = du[i];
            du[i] = d[i + 1];
            d[i + 1] = t - f * d[i + 1];
            if (i + 2 < m) {
                du2[i] = du[i + 1];
                du[i + 1] = -f * du[i + 1];
            }
            piv[i] = 1;
        }
    }
    if (d[m - 1] == 0.0) d[m - 1] = tiny;
    // forward: apply row swaps and L
    for (int i = 0; i + 1 < m; ++i) {
        if (piv[i]) std::swap(x[i], x[i + 1]);
        x[i + 1] -= dl[i] * x[i];
    }
    // backward with U (d, du, du2)
    x[m - 1] /= d[m - 1];
    if (m > 1) x[m - 2] = (x[m - 2] - du[m - 2] * x[m - 1]) / d[m - 2];
    for (int i = m - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
}

void unit(std::vector<double> &x) {
    double s = 0.0;
    for (double v : x) s += v * v;
    s = std::sqrt(s);
    if (s > 0)
        for (double &v : x) v /= s;
}


// Unit eigenvector of T for the eigenvalue lam (three inverse-iteration steps
// from a fixed, slightly varied start; seed varies the start between
// eigenvalues), with its first non-negligible component positive.
void eigvec_at(int m, const double *a, const double *b, double lam, double tiny, int seed, std::vector<double> &x) {
    x.assign(m, 0.0);
    for (int i = 0; i < m; ++i) x[i] = 1.0 + 1e-3 * (double)((i * 7919 + seed * 104729) % 97) / 97.0;
    for (int rep = 0; rep < 3; ++rep) {
        shifted_solve(m, a, b, lam, tiny, x);
        unit(x);
    }
    int lead = 0;
    while (lead + 1 < m && std::fabs(x[lead]) < 1e-300) ++lead;
    if (x[lead] < 0)
        for (double &v : x) v = -v;
}

double tridiag_scale(int m, const double *a, const double *b) {
    double lo = a[0], hi = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(b[i]) : 0.0) + (i + 1 < m ? std::fabs(b[i + 1]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    return std::max({std::fabs(lo), std::fabs(hi), 1e-300});
}

}  // namespace

void tridiag_vector(int m, const double *a, const double *b, double lam, double *z) {
    if (m <= 0) return;
    std::vector<double> x;
    eigvec_at(m, a, b, lam, DBL_EPSILON * tridiag_scale(m, a, b), 0, x);
    for (int i = 0; i < m; ++i) z[i] = x[i];
}

void tridiag_highest(int m, const double *a, const double *b, int k, double *w, double *zlast) {
    if (m <= 0 || k <= 0) return;
    // the lowest of [-a, b], which has the spectrum of -T (and the eigenvectors of
    // T up to alternating signs, so the same |last component|)
    std::vector<double> na(m);
    for (int i = 0; i < m; ++i) na[i] = -a[i];
    tridiag_lowest(m, na.data(), b, k, w, zlast, nullptr);
    for (int j = 0; j < std::min(k, m); ++j) w[j] = -w[j];
}

void tridiag_lowest(int m, const double *a, const double *b, int k, double *w, double *zlast, double *z0) {
    if (m <= 0 || k <= 0) return;
    k = std::min(k, m);
    double bmax = 0.0, lo = a[0], hi = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(b[i]) : 0.0) + (i + 1 < m ? std::fabs(b[i + 1]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
        if (i > 0) bmax = std::max(bmax, std::fabs(b[i]));
    }
    const double scale = std::max({std::fabs(lo), std::fabs(hi), 1e-300});
    const double pivmin = DBL_MIN * std::max(1.0, bmax * bmax);
    lo -= 2.0 * DBL_EPSILON * scale + pivmin;
    hi += 2.0 * DBL_EPSILON * scale + pivmin;
    for (int j = 0; j < k; ++j) w[j] = bisect(m, a, b, j, lo, hi, pivmin);
    const double tiny = DBL_EPSILON * scale;
    std::vector<double> x;
    for (int j = 0; j < k; ++j) {
        eigvec_at(m, a, b, w[j], tiny, j, x);
        zlast[j] = x[m - 1];
        if (j == 0 && z0)
            for (int i = 0; i < m; ++i) z0[i] = x[i];
    }
}

}  // namespace qmb
