// Eigenvalues of a real symmetric matrix (host code): Householder reduction to
// tridiagonal form followed by the implicit QL iteration with Wilkinson shifts
// (Golub & Van Loan, Matrix Computations, 4th ed., sections 8.3.1 and 8.3.5).
// Used for the small D x D problems of the critical-temperature search; plain
// C++ with no device code, so the tests can build it on its own.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

extern "C" int dg_sym_eigvals(int n, const double *A, long lda, double *w);

namespace {

// A (n x n, row-major, full storage) -> diagonal d and sub-diagonal e
// (e[i] couples i and i + 1; e[n-1] = 0).  A is overwritten.
void tridiagonalize(int n, std::vector<double> &A, std::vector<double> &d,
                    std::vector<double> &e) {
    std::vector<double> v(n), p(n);
    for (int k = 0; k + 2 < n; ++k) {
        // Householder vector annihilating A[k+2.., k]
        double sigma = 0.0;
        for (int i = k + 1; i < n; ++i) sigma += A[(size_t)i * n + k] * A[(size_t)i * n + k];
        const double x0 = A[(size_t)(k + 1) * n + k];
        const double norm = std::sqrt(sigma);
        if (norm == 0.0) continue;
        const double alpha = (x0 > 0.0) ? -norm : norm;
        for (int i = k + 1; i < n; ++i) v[i] = A[(size_t)i * n + k];
        v[k + 1] -= alpha;
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        // A_sub <- H A_sub H, H = I - 2 v v^T / (v^T v): with p = 2 A v / vv and
        // q = p - (v^T p / vv) v, A_sub -= v q^T + q v^T
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * v[j];
            p[i] = 2.0 * s / vv;
            vp += v[i] * p[i];
        }
        const double c = vp / vv;
        for (int i = k + 1; i < n; ++i) p[i] -= c * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j)
                A[(size_t)i * n + j] -= v[i] * p[j] + p[i] * v[j];
        A[(size_t)(k + 1) * n + k] = alpha;
        A[(size_t)k * n + (k + 1)] = alpha;
        for (int i = k + 2; i < n; ++i) {
            A[(size_t)i * n + k] = 0.0;
            A[(size_t)k * n + i] = 0.0;
        }
    }
    for (int i = 0; i < n; ++i) {
        d[i] = A[(size_t)i * n + i];
        e[i] = (i + 1 < n) ? A[(size_t)(i + 1) * n + i] : 0.0;
    }
}

// Implicit QL on the symmetric tridiagonal (d, e); eigenvalues left in d.
// An off-diagonal element is negligible below machine epsilon times its two
// diagonal neighbours or times the norm of the whole matrix; the second test
// deflates clusters of (near-)zero eigenvalues, as in rank-deficient scatter
// matrices, and perturbs eigenvalues by at most eps * |T| (the usual
// backward-error bound).  Returns false if an eigenvalue does not converge in
// 64 sweeps.
bool ql_implicit(int n, std::vector<double> &d, std::vector<double> &e) {
    const double eps = std::numeric_limits<double>::epsilon();
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::fabs(d[i]) + std::fabs(e[i]) + (i ? std::fabs(e[i - 1]) : 0.0));
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        while (true) {
            // the first negligible sub-diagonal element at or after l
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                const double em = std::fabs(e[m]);
                if (em <= eps * dd || em <= eps * tnorm) break;
            }
            if (m == l) break;
            if (++iter > 64) return false;
            // Wilkinson shift from the leading 2 x 2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {          // split: restart on the shorter block
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

}  // namespace

// Eigenvalues (unordered) of the symmetric n x n matrix A (lower triangle
// used, row stride lda) into w.  Returns 0, 1 if the iteration failed, or 2
// if its workspace could not be allocated.
extern "C" int dg_sym_eigvals(int n, const double *A, long lda, double *w) {
    if (n <= 0) return 0;
    std::vector<double> M, d, e;
    try {
        M.resize((size_t)n * n);
        d.resize(n);
        e.resize(n);
    } catch (...) {
        return 2;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const double a = A[(size_t)i * lda + j];
            M[(size_t)i * n + j] = a;
            M[(size_t)j * n + i] = a;
        }
    tridiagonalize(n, M, d, e);
    if (!ql_implicit(n, d, e)) return 1;
    for (int i = 0; i < n; ++i) w[i] = d[i];
    return 0;
}
