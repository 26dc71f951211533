#include "dg_init.hpp"

#include <cfloat>
#include <cmath>
#include <ctime>

namespace dgh {

MT19937::MT19937(uint32_t seed) {
    if (seed == 0) seed = 4357;
    mt_[0] = seed;
    for (int i = 1; i < 624; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
    mti_ = 624;
}

uint32_t MT19937::next() {
    if (mti_ >= 624) {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[(i + 1) % 624] & 0x7fffffffu);
            mt_[i] = mt_[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        mti_ = 0;
    }
    uint32_t y = mt_[mti_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

Mat rctg(int g, int L) {
    long n = 1;
    for (int l = 0; l < L; ++l) n *= g;
    Mat X(n, L);
    const double sc = 2.0 / (double)(g - 1);
    for (long i = 0; i < n; ++i) {
        long r = i;
        for (int l = 0; l < L; ++l) {       // coordinate 0 varies fastest
            X(i, l) = (double)(r % g) * sc + (-1.0);
            r /= g;
        }
    }
    return X;
}

static long ipow(long b, int e) {
    long r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

bool setup_basis(int L, int K, int M, double s, Mat &X, Mat &FI) {
    if (L < 1 || L > 3 || s <= 0) return false;
    const int gx = (int)std::lround(std::pow((double)K, 1.0 / L));
    const int gm = (int)std::lround(std::pow((double)M, 1.0 / L));
    if (ipow(gx, L) != K || ipow(gm, L) != M || gx < 2 || gm < 2) return false;
    X = rctg(gx, L);
    Mat MU = rctg(gm, L);
    const double scale = (double)gm / (gm - 1);
    for (double &v : MU.v) v *= scale;
    const double sigma = s * (MU(1, 0) - MU(0, 0));
    const double c = -0.5 / (sigma * sigma);
    FI = Mat(K, M + 1);
    for (int k = 0; k < K; ++k) {
        for (int m = 0; m < M; ++m) {
            double d = 0.0;
            for (int l = 0; l < L; ++l) {
                const double t = X(k, l) - MU(m, l);
                d += t * t;
            }
            FI(k, m) = std::exp(d * c);
        }
        FI(k, M) = 1.0;
    }
    return true;
}

void mean_sd(const Mat &T, std::vector<double> &mu, std::vector<double> &sd) {
    const long N = T.rows, D = T.cols;
    mu.assign(D, 0.0);
    sd.assign(D, 0.0);
    for (long i = 0; i < N; ++i)
        for (long j = 0; j < D; ++j) mu[j] += T(i, j);
    for (long j = 0; j < D; ++j) mu[j] *= 1.0 / N;
    for (long i = 0; i < N; ++i)
        for (long j = 0; j < D; ++j) {
            const double c = T(i, j) - mu[j];
            sd[j] += c * c;
        }
    for (long j = 0; j < D; ++j) sd[j] = std::sqrt(sd[j] * (1.0 / (N - 1)));
}

Mat random_centres(const std::vector<double> &mu, const std::vector<double> &sd, int K,
                   unsigned long seed) {
    if (seed == 0) seed = (unsigned long)std::time(nullptr);
    MT19937 r((uint32_t)seed);
    const long D = (long)mu.size();
    Mat Y(K, D);
    for (int i = 0; i < K; ++i)
        for (long j = 0; j < D; ++j) {
            const double dev = sd[j] * 6;
            const double u = std::ceil(r.uniform() / 1e-15) * 1e-15;
            Y(i, j) = 2 * dev * u - dev + mu[j];
        }
    return Y;
}

void eig_sym(int n, std::vector<double> &A, std::vector<double> &ev, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    auto v = [&](int i, int j) -> double & { return V[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += a(i, i) * a(i, i);
            for (int j = i + 1; j < n; ++j) off += a(i, j) * a(i, j);
        }
        if (off == 0.0 || off <= 1e-32 * diag) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a(p, q);
                if (apq == 0.0) continue;
                const double th = (a(q, q) - a(p, p)) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int r = 0; r < n; ++r) {
                    const double x = a(r, p), y = a(r, q);
                    a(r, p) = cs * x - sn * y;
                    a(r, q) = sn * x + cs * y;
                }
                for (int r = 0; r < n; ++r) {
                    const double x = a(p, r), y = a(q, r);
                    a(p, r) = cs * x - sn * y;
                    a(q, r) = sn * x + cs * y;
                }
                a(p, q) = a(q, p) = 0.0;
                for (int r = 0; r < n; ++r) {
                    const double x = v(r, p), y = v(r, q);
                    v(r, p) = cs * x - sn * y;
                    v(r, q) = sn * x + cs * y;
                }
            }
    }
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = a(i, i);
}

Mat least_squares(const Mat &FI, const Mat &Y) {
    const long K = FI.rows, n = FI.cols, D = Y.cols;
    std::vector<double> A((size_t)n * n, 0.0);
    Mat B(n, D);
    for (long k = 0; k < K; ++k)
        for (long i = 0; i < n; ++i) {
            const double f = FI(k, i);
            for (long j = 0; j < n; ++j) A[(size_t)i * n + j] += f * FI(k, j);
            for (long d = 0; d < D; ++d) B(i, d) += f * Y(k, d);
        }
    std::vector<double> ev, V;
    eig_sym((int)n, A, ev, V);
    Mat W(n, D);
    std::vector<double> w(n);
    for (long c = 0; c < D; ++c) {
        for (long i = 0; i < n; ++i) {
            double s = 0.0;
            for (long r = 0; r < n; ++r) s += V[(size_t)r * n + i] * B(r, c);
            w[i] = (ev[i] != 0.0) ? s / ev[i] : 0.0;
        }
        for (long r = 0; r < n; ++r) {
            double s = 0.0;
            for (long i = 0; i < n; ++i) s += V[(size_t)r * n + i] * w[i];
            W(r, c) = s;
        }
    }
    return W;
}

double initial_beta(const Mat &Y) {
    const long K = Y.rows, D = Y.cols;
    long double mean = 0;
    for (long j = 0; j < K; ++j) {
        double mn = DBL_MAX;
        for (long i = 0; i < K; ++i) {
            if (i == j) continue;
            double d = 0.0;
            for (long l = 0; l < D; ++l) {
                const double t = Y(i, l) - Y(j, l);
                d += t * t;
            }
            if (d < mn) mn = d;
        }
        mean += (mn - mean) / (long double)(j + 1);
    }
    double m = (double)mean;
    if (m == 0) m = 1;
    return 2 / m;
}

double first_critical_temp(const Mat &T) {
    const long N = T.rows, D = T.cols;
    std::vector<double> c(D, 0.0);
    for (long i = 0; i < N; ++i)
        for (long j = 0; j < D; ++j) c[j] += T(i, j);
    for (long j = 0; j < D; ++j) c[j] *= 1.0 / N;
    std::vector<double> S((size_t)D * D, 0.0), row(D);
    double ss = 0.0;
    for (long i = 0; i < N; ++i) {
        for (long j = 0; j < D; ++j) {
            row[j] = T(i, j) - c[j];
            ss += row[j] * row[j];
        }
        for (long a = 0; a < D; ++a) {
            const double ra = row[a];
            if (ra == 0.0) continue;
            double *s = &S[(size_t)a * D];
            for (long b = 0; b < D; ++b) s[b] += ra * row[b];
        }
    }
    std::vector<double> ev, V;
    eig_sym((int)D, S, ev, V);
    double lmax = ev[0];
    for (double e : ev) lmax = e > lmax ? e : lmax;
    lmax /= N;
    const double betaT = (double)N * D / ss;
    return lmax * betaT;
}

}  // namespace dgh
