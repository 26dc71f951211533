// Host-side model set-up of the dagtm driver: latent grid, RBF basis, the
// seeded random start, least-squares W0, beta0 and the first critical
// temperature.  Restates the reference's fresh-start chain
// (dagtm.cxx:1104-1159; comm.c:992-1185, 1247-1294, 1401-1536;
// mpicomm.c:511-559, 686-730) in FP64 on the CPU; it runs once per run, on
// host_threads() workers.  Every output element is computed by one worker in
// a fixed order, so the results do not depend on the number of threads.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace dgh {

// Workers for the set-up chain: DG_HOST_THREADS when set, else the CPUs this
// process may run on, at most 16.
int host_threads();

// A fixed set of worker threads; run(f) calls f(t) for every worker t in
// [0, size()) (the calling thread is worker 0) and returns when all are done.
class Pool {
   public:
    explicit Pool(int n);
    ~Pool();
    Pool(const Pool &) = delete;
    Pool &operator=(const Pool &) = delete;
    int size() const { return n_; }
    void run(const std::function<void(int)> &f);

   private:
    void loop(int t);
    int n_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    const std::function<void(int)> *f_ = nullptr;
    std::atomic<unsigned long> gen_{0};
    std::atomic<int> left_{0};
    std::atomic<bool> stop_{false};
};

// The process's pool of host_threads() workers, made on first use.
Pool &pool();

// Row-major dense matrix of doubles.
struct Mat {
    long rows = 0, cols = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(long r, long c) : rows(r), cols(c), v((size_t)r * c, 0.0) {}
    double &operator()(long i, long j) { return v[(size_t)i * cols + j]; }
    const double &operator()(long i, long j) const { return v[(size_t)i * cols + j]; }
    double *data() { return v.data(); }
    const double *data() const { return v.data(); }
};

// MT19937 as GSL's default generator: init_genrand seeding, seed 0 -> 4357,
// uniform = u / 2^32.
class MT19937 {
   public:
    explicit MT19937(uint32_t seed);
    uint32_t next();
    double uniform() { return next() / 4294967296.0; }

   private:
    uint32_t mt_[624];
    int mti_;
};

// g^L grid on [-1,1]^L, first coordinate fastest (comm.c:1453-1504).
Mat rctg(int g, int L);
// X (K x L) and FI (K x (M+1)) for the sweep grid, width sigma = s * spacing
// (comm.c:992-1103, 1506-1536).  Returns false if K or M is not a perfect
// L-th power.
bool setup_basis(int L, int K, int M, double s, Mat &X, Mat &FI);
// Column mean and sample standard deviation (mpicomm.c:686-730).
void mean_sd(const Mat &T, std::vector<double> &mu, std::vector<double> &sd);
// Uniform random centres on mu +- 6 sd (comm.c:1401-1448).
Mat random_centres(const std::vector<double> &mu, const std::vector<double> &sd, int K,
                   unsigned long seed);
// Least-squares W minimising |FI W - Y| through the normal equations and a
// symmetric eigen pseudo-inverse that drops exact-zero eigenvalues
// (comm.c:1247-1294).  Empty if the eigen-decomposition does not converge.
Mat least_squares(const Mat &FI, const Mat &Y);
// A B, each element summed over the inner index in order.
Mat product(const Mat &A, const Mat &B);
// 2 / mean nearest-neighbour squared distance (comm.c:1161-1185).
double initial_beta(const Mat &Y);
// First critical temperature lambda_max(Cov T) * beta_T (mpicomm.c:511-559).
double first_critical_temp(const Mat &T);
// Eigenvalues ev (unordered) of the symmetric n x n row-major A (overwritten)
// by Householder tridiagonalisation and implicit QL with Wilkinson shifts;
// with Vt, eigenvector i is row i of Vt (n x n).  False if QL does not
// converge.
bool eig_sym(int n, std::vector<double> &A, std::vector<double> &ev, std::vector<double> *Vt);

}  // namespace dgh
