#include "dg_init.hpp"

#include <sched.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <ctime>
#include <cstring>
#include <limits>
#include <memory>
#include <new>

namespace dgh {

// ---- host threads

int host_threads() {
    if (const char *s = std::getenv("DG_HOST_THREADS")) {
        const int n = std::atoi(s);
        if (n >= 1) return std::min(n, 256);
    }
    cpu_set_t set;
    int n = 0;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (n < 1) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 16));
}

Pool::Pool(int n) : n_(std::max(1, n)) {
    for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
}

Pool::~Pool() {
    stop_.store(true);
    gen_.fetch_add(1);
    {
        std::lock_guard<std::mutex> g(m_);
    }
    go_.notify_all();
    for (std::thread &t : th_) t.join();
}

namespace {

// Spins up to about 200 microseconds for done() before the caller blocks:
// the pool's calls come back to back inside one set-up step, and a sleeping
// worker can take longer than that to wake.
template <class C>
bool spin_until(const C &done) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0;; ++i) {
        if (done()) return true;
        if ((i & 63) == 63 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200))
            return false;
        std::this_thread::yield();
    }
}

}  // namespace

void Pool::run(const std::function<void(int)> &f) {
    if (n_ == 1) {
        f(0);
        return;
    }
    f_ = &f;
    left_.store(n_ - 1);
    gen_.fetch_add(1);
    {
        std::lock_guard<std::mutex> g(m_);   // orders the change before a waiter's test
    }
    go_.notify_all();
    f(0);
    auto finished = [this] { return left_.load() == 0; };
    if (spin_until(finished)) return;
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, finished);
}

void Pool::loop(int t) {
    unsigned long seen = 0;
    for (;;) {
        auto started = [&] { return gen_.load() != seen; };
        if (!spin_until(started)) {
            std::unique_lock<std::mutex> g(m_);
            go_.wait(g, started);
        }
        seen = gen_.load();
        if (stop_.load()) return;
        (*f_)(t);
        if (left_.fetch_sub(1) == 1) {
            {
                std::lock_guard<std::mutex> g(m_);
            }
            done_.notify_one();
        }
    }
}

Pool &pool() {
    static Pool p(host_threads());
    return p;
}

namespace {

// [b, e) of n items for worker t of nt, in contiguous blocks.
inline void block(long n, int t, int nt, long &b, long &e) {
    b = n * t / nt;
    e = n * (t + 1) / nt;
}

// f(i) for i in [0, n), in contiguous blocks over the pool; serial when the
// whole loop costs less than about `work` multiply-adds.
template <class F>
void par_rows(long n, double work, const F &f) {
    Pool &P = pool();
    if (P.size() == 1 || work < 2e5 || n < 2) {
        for (long i = 0; i < n; ++i) f(i);
        return;
    }
    P.run([&](int t) {
        long b, e;
        block(n, t, P.size(), b, e);
        for (long i = b; i < e; ++i) f(i);
    });
}

// Rows of an n x n triangle (row i costing n - i) for worker t of nt: the
// rows i with i mod 2nt equal to t or 2nt - 1 - t, so every worker gets about
// the same work.
template <class F>
void tri_rows(long n, int t, int nt, const F &f) {
    for (long i = t; i < n; i += 2 * nt) f(i);
    for (long i = 2 * nt - 1 - t; i < n; i += 2 * nt) f(i);
}

// Zero-filled row-major matrix whose rows start on 64-byte cache lines (the
// row stride is padded to whole lines), so workers that write different rows,
// or different 8-column blocks of one row, never write the same line.
struct LineMat {
    long rows = 0, cols = 0, ld = 0;
    std::unique_ptr<double, void (*)(void *)> p{nullptr, std::free};
    LineMat(long r, long c) : rows(r), cols(c), ld((c + 7) & ~7L) {
        const size_t bytes = std::max<size_t>(64, sizeof(double) * (size_t)r * ld);
        p.reset(static_cast<double *>(std::aligned_alloc(64, bytes)));
        if (!p) throw std::bad_alloc();
        std::memset(p.get(), 0, bytes);
    }
    double *row(long i) { return p.get() + i * ld; }
    const double *row(long i) const { return p.get() + i * ld; }
    double &operator()(long i, long j) { return p.get()[i * ld + j]; }
};

// Column blocks [b, e) for worker t of nt with both ends on multiples of 8.
inline void line_block(long n, int t, int nt, long &b, long &e) {
    const long lines = (n + 7) / 8;
    b = std::min(n, lines * t / nt * 8);
    e = std::min(n, lines * (t + 1) / nt * 8);
}

}  // namespace

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
    par_rows(K, 20.0 * K * M, [&](long k) {
        for (int m = 0; m < M; ++m) {
            double d = 0.0;
            for (int l = 0; l < L; ++l) {
                const double t = X(k, l) - MU(m, l);
                d += t * t;
            }
            FI(k, m) = std::exp(d * c);
        }
        FI(k, M) = 1.0;
    });
    return true;
}

namespace {

// f(j0, j1) over blocks of the D columns of T, one block per worker: each
// worker reads only its columns of every row, so T is read once in all.
template <class F>
void column_blocks(const Mat &T, const F &f) {
    Pool &P = pool();
    const long D = T.cols;
    if (P.size() == 1 || (double)T.rows * D < 2e5 || D < 2) {
        f(0L, D);
        return;
    }
    const int nt = (int)std::min<long>(P.size(), D);
    P.run([&](int t) {
        if (t >= nt) return;
        long b, e;
        block(D, t, nt, b, e);
        f(b, e);
    });
}

// c[j] += sum_i T(i, j), each column summed over i in order.
void column_sums(const Mat &T, std::vector<double> &c) {
    column_blocks(T, [&](long j0, long j1) {
        std::vector<double> acc(c.begin() + j0, c.begin() + j1);
        for (long i = 0; i < T.rows; ++i) {
            const double *x = &T(i, j0);
            for (long j = 0; j < j1 - j0; ++j) acc[j] += x[j];
        }
        std::copy(acc.begin(), acc.end(), c.begin() + j0);
    });
}

}  // namespace

void mean_sd(const Mat &T, std::vector<double> &mu, std::vector<double> &sd) {
    const long N = T.rows, D = T.cols;
    mu.assign(D, 0.0);
    sd.assign(D, 0.0);
    column_sums(T, mu);
    for (long j = 0; j < D; ++j) mu[j] *= 1.0 / N;
    column_blocks(T, [&](long j0, long j1) {
        std::vector<double> acc(j1 - j0, 0.0);
        for (long i = 0; i < N; ++i) {
            const double *x = &T(i, j0);
            for (long j = 0; j < j1 - j0; ++j) {
                const double c = x[j] - mu[j0 + j];
                acc[j] += c * c;
            }
        }
        std::copy(acc.begin(), acc.end(), sd.begin() + j0);
    });
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

namespace {

// ---- symmetric eigenproblem: Householder tridiagonalisation, then implicit
// QL with Wilkinson shifts (Golub & Van Loan, Matrix Computations, 4th ed.,
// 8.3.1 and 8.3.5).  Every element is computed by one worker in a fixed order,
// so the results do not depend on the number of threads.

// A (n x n, full storage) -> diagonal d and sub-diagonal e (e[i] couples i
// and i + 1).  With vectors, Qt (n x n) receives Q^T for A = Q T Q^T.
void tridiagonalize(int n, LineMat &A, std::vector<double> &d, std::vector<double> &e,
                    LineMat *Qt) {
    auto a = [&](long i, long j) -> double & { return A(i, j); };
    std::vector<double> tau(n, 0.0), p(n);
    LineMat U(Qt ? n : 0, n);
    for (int k = 0; k + 2 < n; ++k) {
        // row k of the trailing block equals its column k (A stays symmetric)
        const double *x = A.row(k);
        double sigma = 0.0;
        for (int i = k + 1; i < n; ++i) sigma += x[i] * x[i];
        const double norm = std::sqrt(sigma);
        if (norm == 0.0) continue;
        const double alpha = (x[k + 1] > 0.0) ? -norm : norm;
        std::vector<double> v(x, x + n);
        v[k + 1] -= alpha;
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        // A_sub <- H A_sub H with H = I - 2 v v^T / vv: p = 2 A v / vv,
        // q = p - (v.p / vv) v, A_sub -= v q^T + q v^T
        const long m = n - k - 1;
        par_rows(m, (double)m * m, [&](long r) {
            const long i = k + 1 + r;
            const double *ai = &a(i, 0);
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += ai[j] * v[j];
            p[i] = 2.0 * s / vv;
        });
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) vp += v[i] * p[i];
        const double c = vp / vv;
        for (int i = k + 1; i < n; ++i) p[i] -= c * v[i];
        par_rows(m, 2.0 * m * m, [&](long r) {
            const long i = k + 1 + r;
            double *ai = &a(i, 0);
            const double vi = v[i], pi = p[i];
            for (int j = k + 1; j < n; ++j) ai[j] -= vi * p[j] + pi * v[j];
        });
        a(k + 1, k) = a(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
        if (Qt) {
            for (int i = k + 1; i < n; ++i) U(k, i) = v[i];
            tau[k] = 2.0 / vv;
        }
    }
    d.resize(n);
    e.resize(n);
    for (int i = 0; i < n; ++i) {
        d[i] = a(i, i);
        e[i] = (i + 1 < n) ? a(i + 1, i) : 0.0;
    }
    if (!Qt) return;
    // Q = H_0 H_1 ... H_{n-3}; Q^T = H_{n-3} ... H_0 built from the right:
    // Qt <- Qt H_k touches columns k+1.. of rows k+1.. only
    LineMat &Z = *Qt;
    for (int i = 0; i < n; ++i) Z(i, i) = 1.0;
    for (int k = n - 3; k >= 0; --k) {
        if (tau[k] == 0.0) continue;
        const double *v = U.row(k);
        const double tk = tau[k];
        const long m = n - k - 1;
        par_rows(m, 2.0 * m * m, [&](long r) {
            double *z = Z.row(k + 1 + r);
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += z[j] * v[j];
            s *= tk;
            for (int j = k + 1; j < n; ++j) z[j] -= s * v[j];
        });
    }
}

struct Rot {
    int i;
    double c, s;
};

// Rows i and i + 1 of Zt (n x n) rotated by each of rots in order; the
// workers take column blocks that start on cache lines.
void apply_rotations(int n, LineMat &Zt, const std::vector<Rot> &rots) {
    Pool &P = pool();
    const int nt = (P.size() > 1 && (double)rots.size() * n > 4e5) ? P.size() : 1;
    auto slice = [&](int t) {
        long b, e;
        line_block(n, t, nt, b, e);
        for (const Rot &r : rots) {
            double *zi = Zt.row(r.i), *zj = Zt.row(r.i + 1);
            for (long k = b; k < e; ++k) {
                const double f = zj[k];
                zj[k] = r.s * zi[k] + r.c * f;
                zi[k] = r.c * zi[k] - r.s * f;
            }
        }
    };
    if (nt == 1)
        slice(0);
    else
        P.run([&](int t) { slice(t); });
}

// Implicit QL on the symmetric tridiagonal (d, e); eigenvalues left in d and,
// with Zt (= Q^T of the tridiagonalisation), eigenvector i left in row i of
// Zt.  An off-diagonal element is negligible below machine epsilon times its
// two diagonal neighbours or times the norm of the whole matrix (the second
// test deflates clusters of near-zero eigenvalues of rank-deficient normal
// equations).  Returns false if an eigenvalue does not converge in 64 sweeps.
bool ql_implicit(int n, std::vector<double> &d, std::vector<double> &e, LineMat *Zt) {
    const double eps = std::numeric_limits<double>::epsilon();
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::fabs(d[i]) + std::fabs(e[i]) + (i ? std::fabs(e[i - 1]) : 0.0));
    std::vector<Rot> rots;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                const double em = std::fabs(e[m]);
                if (em <= eps * dd || em <= eps * tnorm) break;
            }
            if (m == l) break;
            if (++iter > 64) return false;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool split = false;
            rots.clear();
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {          // split: restart on the shorter block
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    split = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (Zt) rots.push_back({i, c, s});
            }
            if (Zt && !rots.empty()) apply_rotations(n, *Zt, rots);
            if (split) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

}  // namespace

namespace {

// Eigen-decomposition of the symmetric A (overwritten); with Vt, eigenvector
// i in row i.
bool eig_lines(int n, LineMat &A, std::vector<double> &ev, LineMat *Vt) {
    std::vector<double> e;
    tridiagonalize(n, A, ev, e, Vt);
    return ql_implicit(n, ev, e, Vt);
}

}  // namespace

bool eig_sym(int n, std::vector<double> &A, std::vector<double> &ev, std::vector<double> *Vt) {
    if (n <= 0) return true;
    LineMat M(n, n), Z(Vt ? n : 0, n);
    for (int i = 0; i < n; ++i) std::copy(&A[(size_t)i * n], &A[(size_t)i * n] + n, M.row(i));
    if (!eig_lines(n, M, ev, Vt ? &Z : nullptr)) return false;
    for (int i = 0; i < n; ++i) std::copy(M.row(i), M.row(i) + n, &A[(size_t)i * n]);
    if (Vt) {
        Vt->resize((size_t)n * n);
        for (int i = 0; i < n; ++i) std::copy(Z.row(i), Z.row(i) + n, &(*Vt)[(size_t)i * n]);
    }
    return true;
}

Mat least_squares(const Mat &FI, const Mat &Y) {
    const long K = FI.rows, n = FI.cols, D = Y.cols;
    // normal equations A = FI^T FI (upper triangle, then mirrored) and
    // B = FI^T Y, every element summed over k in order; the rows of FI go
    // through in panels so that the workers share them in cache
    LineMat A(n, n), B(n, D);
    Pool &P = pool();
    const long panel = 64;
    for (long k0 = 0; k0 < K; k0 += panel) {
        const long k1 = std::min(K, k0 + panel);
        auto rows = [&](int t, int nt) {
            tri_rows(n, t, nt, [&](long i) {
                double *ai = A.row(i);
                double *bi = B.row(i);
                long k = k0;
                // four rows of FI per pass over the output row; the adds of
                // every element still go in the order of k
                for (; k + 4 <= k1; k += 4) {
                    const double *f0 = &FI(k, 0), *f1 = f0 + n, *f2 = f1 + n, *f3 = f2 + n;
                    const double g0 = f0[i], g1 = f1[i], g2 = f2[i], g3 = f3[i];
                    for (long j = i; j < n; ++j) {
                        double s = ai[j];
                        s += g0 * f0[j];
                        s += g1 * f1[j];
                        s += g2 * f2[j];
                        s += g3 * f3[j];
                        ai[j] = s;
                    }
                    const double *y0 = &Y(k, 0), *y1 = y0 + D, *y2 = y1 + D, *y3 = y2 + D;
                    for (long d = 0; d < D; ++d) {
                        double s = bi[d];
                        s += g0 * y0[d];
                        s += g1 * y1[d];
                        s += g2 * y2[d];
                        s += g3 * y3[d];
                        bi[d] = s;
                    }
                }
                for (; k < k1; ++k) {
                    const double *f = &FI(k, 0);
                    const double fi = f[i];
                    for (long j = i; j < n; ++j) ai[j] += fi * f[j];
                    const double *y = &Y(k, 0);
                    for (long d = 0; d < D; ++d) bi[d] += fi * y[d];
                }
            });
        };
        if (P.size() == 1)
            rows(0, 1);
        else
            P.run([&](int t) { rows(t, P.size()); });
    }
    for (long i = 0; i < n; ++i)
        for (long j = 0; j < i; ++j) A(i, j) = A(j, i);
    // symmetric eigen pseudo-inverse: W = V diag(1/ev) V^T B, eigenvalues
    // that are exactly zero dropped
    std::vector<double> ev;
    LineMat Zt(n, n);
    if (!eig_lines((int)n, A, ev, &Zt)) return Mat();
    LineMat C(n, D), Wl(n, D);
    par_rows(n, (double)n * n * D, [&](long i) {
        const double *z = Zt.row(i);
        double *ci = C.row(i);
        for (long r = 0; r < n; ++r) {
            const double zr = z[r];
            const double *br = B.row(r);
            for (long d = 0; d < D; ++d) ci[d] += zr * br[d];
        }
        const double inv = (ev[i] != 0.0) ? 1.0 / ev[i] : 0.0;
        for (long d = 0; d < D; ++d) ci[d] = (ev[i] != 0.0) ? ci[d] * inv : 0.0;
    });
    par_rows(n, (double)n * n * D, [&](long r) {
        double *wr = Wl.row(r);
        for (long i = 0; i < n; ++i) {
            const double z = Zt(i, r);
            const double *ci = C.row(i);
            for (long d = 0; d < D; ++d) wr[d] += z * ci[d];
        }
    });
    Mat W(n, D);
    for (long r = 0; r < n; ++r) std::copy(Wl.row(r), Wl.row(r) + D, &W(r, 0));
    return W;
}

Mat product(const Mat &A, const Mat &B) {
    const long R = A.rows, I = A.cols, Cc = B.cols;
    Mat out(R, Cc);
    par_rows(R, (double)R * I * Cc, [&](long r) {
        const double *ar = &A(r, 0);
        double *o = &out(r, 0);          // starts at zero; adds in the order of i
        for (long i = 0; i < I; ++i) {
            const double a = ar[i];
            const double *b = &B(i, 0);
            for (long c = 0; c < Cc; ++c) o[c] += a * b[c];
        }
    });
    return out;
}

double initial_beta(const Mat &Y) {
    const long K = Y.rows, D = Y.cols;
    std::vector<double> mn(K, DBL_MAX);
    par_rows(K, (double)K * K * D, [&](long j) {
        const double *yj = &Y(j, 0);
        double best = DBL_MAX;
        for (long i = 0; i < K; ++i) {
            if (i == j) continue;
            const double *yi = &Y(i, 0);
            double d = 0.0;
            for (long l = 0; l < D; ++l) {
                const double t = yi[l] - yj[l];
                d += t * t;
            }
            if (d < best) best = d;
        }
        mn[j] = best;
    });
    long double mean = 0;
    for (long j = 0; j < K; ++j) mean += (mn[j] - mean) / (long double)(j + 1);
    double m = (double)mean;
    if (m == 0) m = 1;
    return 2 / m;
}

double first_critical_temp(const Mat &T) {
    const long N = T.rows, D = T.cols;
    std::vector<double> c(D, 0.0);
    column_sums(T, c);
    for (long j = 0; j < D; ++j) c[j] *= 1.0 / N;
    // scatter S = sum_i (t_i - c)(t_i - c)^T, upper triangle, every element
    // summed over i in order; rows are centred a panel at a time into a
    // buffer the workers share
    LineMat S(D, D);
    const long panel = 1024;
    LineMat buf(panel, D);
    Pool &P = pool();
    for (long i0 = 0; i0 < N; i0 += panel) {
        const long np = std::min(N, i0 + panel) - i0;
        auto centre = [&](int t, int nt) {
            long b, e;
            block(np, t, nt, b, e);
            for (long r = b; r < e; ++r) {
                const double *x = &T(i0 + r, 0);
                double *y = buf.row(r);
                for (long j = 0; j < D; ++j) y[j] = x[j] - c[j];
            }
        };
        auto scatter = [&](int t, int nt) {
            tri_rows(D, t, nt, [&](long a) {
                double *s = S.row(a);
                long r = 0;
                for (; r + 4 <= np; r += 4) {    // element adds in the order of r
                    const double *y0 = buf.row(r), *y1 = buf.row(r + 1), *y2 = buf.row(r + 2),
                                 *y3 = buf.row(r + 3);
                    const double a0 = y0[a], a1 = y1[a], a2 = y2[a], a3 = y3[a];
                    for (long b = a; b < D; ++b) {
                        double v = s[b];
                        v += a0 * y0[b];
                        v += a1 * y1[b];
                        v += a2 * y2[b];
                        v += a3 * y3[b];
                        s[b] = v;
                    }
                }
                for (; r < np; ++r) {
                    const double *y = buf.row(r);
                    const double ya = y[a];
                    for (long b = a; b < D; ++b) s[b] += ya * y[b];
                }
            });
        };
        if (P.size() == 1) {
            centre(0, 1);
            scatter(0, 1);
        } else {
            P.run([&](int t) { centre(t, P.size()); });
            P.run([&](int t) { scatter(t, P.size()); });
        }
    }
    double ss = 0.0;
    for (long a = 0; a < D; ++a) ss += S(a, a);
    for (long a = 0; a < D; ++a)
        for (long b = 0; b < a; ++b) S(a, b) = S(b, a);
    std::vector<double> ev;
    if (!eig_lines((int)D, S, ev, nullptr)) return std::nan("");
    double lmax = ev[0];
    for (double x : ev) lmax = x > lmax ? x : lmax;
    lmax /= N;
    const double betaT = (double)N * D / ss;
    return lmax * betaT;
}

}  // namespace dgh
