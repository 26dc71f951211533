// FP64 M-step and the scalar reductions of DA-GTM on gfx950.
//
//   A  = FI^T diag(G) FI  (scale_by_row, then dgemm)  dagtm.cxx:1729-1738
//   B  = FI^T (R T) = FI^T RT                dagtm.cxx:1746-1761 (same product,
//                                            associated the other way)
//   W  = (A + lambda/beta I)^-1 B            Cholesky (dagtm_solve2 in the
//                                            reference, comm.c:1299-1348)
//   Y  = FI W                                dagtm.cxx:1769-1770
//   beta = N D / sum_kn R_kn |y_k - t_n|^2   mpicomm.c:276-302, evaluated as
//          sum_n |t_n|^2 - 2 sum_k y_k.RT_k + sum_k G_k |y_k|^2
//          (exact because every column of R sums to one, mpicomm.c:382-391)
#include <cmath>

#include "dg_internal.h"

namespace {

// out[r] = sum_d X[r][d]^2, sequential in d (the ddot loops of comm.c:604-618)
__device__ __forceinline__ void k_rownorm_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ X,
    long rows, int ld, int D, double *__restrict__ out) {
    const long r = (long)bx * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    // eight loads in flight, then their FMAs in order
    double s = 0.0;
    int d = 0;
    for (; d + 8 <= D; d += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = X[r * ld + d + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) s = fma(v[u], v[u], s);
    }
    for (; d < D; ++d) {
        const double v = X[r * ld + d];
        s = fma(v, v, s);
    }
    out[r] = s;
}

__global__ void k_rownorm(const double *__restrict__ X, long rows, int ld, int D,
                          double *__restrict__ out) {
    k_rownorm_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, X, rows, ld, D, out);
}

// Deterministic block-tree sum of one or two vectors into scal[slot], scal[slot+1].
// Single workgroup, FP64.
__global__ __launch_bounds__(1024) void k_sum2(const double *__restrict__ a,
                                               const double *__restrict__ b, long n,
                                               double *__restrict__ scal, int slot) {
    __shared__ double sa[1024], sb[1024];
    double x = 0.0, y = 0.0;
    for (long i = threadIdx.x; i < n; i += 1024) {
        x += a[i];
        if (b) y += b[i];
    }
    sa[threadIdx.x] = x; sb[threadIdx.x] = y;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[slot] = sa[0];
        if (b) scal[slot + 1] = sb[0];
    }
}

// Number of zero columns (pm == -inf) into scal[kScalNZero].
__global__ __launch_bounds__(1024) void k_count_zero(const double *__restrict__ pm, long n,
                                                     double *__restrict__ scal) {
    __shared__ double s[1024];
    double c = 0.0;
    for (long i = threadIdx.x; i < n; i += 1024) c += (pm[i] == -INFINITY) ? 1.0 : 0.0;
    s[threadIdx.x] = c;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) scal[kScalNZero] = s[0];
}

// tr(A0) into *out, summed by the first wave; every thread of the workgroup
// calls it (it ends with a barrier).  Only the jitter retries need it.
__device__ void block_trace(const double *__restrict__ A0, int n, double *out) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        double s = 0.0;
        for (int i = tid; i < n; i += 64) s += A0[(long)i * n + i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (tid == 0) *out = s;
    }
    __syncthreads();
}

// Cholesky solve of (A0 + (ridge + jitter) I) W = B in one workgroup (the
// batched M-step's path for M + 1 > 126; a single instance uses the blocked
// kernels below), the factor built in place in A (right-looking; each entry
// receives its updates in the same column order as a left-looking loop).  On a
// non-positive pivot the factorisation restarts with jitter 1e-13 tr/n,
// growing x10 per retry, at most 8 retries.  scal[kScalSolveOK] = 1 on
// success, scal[kScalJitter] = the jitter used.
__device__ __forceinline__ void k_chol_solve_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ A0,
    double *__restrict__ A, int n, double ridge, const double *__restrict__ B, int nrhs,
    double *__restrict__ Wout, double *__restrict__ scal) {
    __shared__ int fail;
    __shared__ double tr;
    const int tid = threadIdx.x, nt = blockDim.x;
    double jit = 0.0;
    int ok = 0;
    for (int attempt = 0; attempt < 9 && !ok; ++attempt) {
        if (attempt == 1) {
            block_trace(A0, n, &tr);
            jit = 1e-13 * tr / n;
        } else if (attempt > 1) {
            jit *= 10.0;
        }
        for (long e = tid; e < (long)n * n; e += nt) {
            const long i = e / n, j = e % n;
            A[e] = A0[e] + ((i == j) ? ridge + jit : 0.0);
        }
        if (tid == 0) fail = 0;
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            if (tid == 0) {
                const double p = A[(long)j * n + j];
                if (!(p > 0.0)) fail = 1;
                else A[(long)j * n + j] = sqrt(p);
            }
            __syncthreads();
            if (fail) break;
            const double ljj = A[(long)j * n + j];
            for (int i = j + 1 + tid; i < n; i += nt) A[(long)i * n + j] /= ljj;
            __syncthreads();
            // trailing update of the lower triangle, columns j+1..n-1
            const int m = n - j - 1;
            const long pairs = (long)m * (m + 1) / 2;
            for (long e = tid; e < pairs; e += nt) {
                // e -> (r, c) with 0 <= c <= r < m
                long r = (long)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
                while ((r + 1) * (r + 2) / 2 <= e) ++r;
                while (r * (r + 1) / 2 > e) --r;
                const long cc = e - r * (r + 1) / 2;
                const long i = j + 1 + r, k = j + 1 + cc;
                A[i * n + k] = fma(-A[i * n + j], A[k * n + j], A[i * n + k]);
            }
            __syncthreads();
        }
        ok = !fail;
        __syncthreads();
    }
    if (tid == 0) {
        scal[kScalJitter] = jit;
        scal[kScalSolveOK] = ok ? 1.0 : 0.0;
    }
    if (!ok) return;
    // forward L y = b and backward L^T x = y, one right-hand side per thread
    for (int c = tid; c < nrhs; c += nt) {
        for (int i = 0; i < n; ++i) {
            double s = B[(long)i * nrhs + c];
            for (int p = 0; p < i; ++p) s = fma(-A[(long)i * n + p], Wout[(long)p * nrhs + c], s);
            Wout[(long)i * nrhs + c] = s / A[(long)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = Wout[(long)i * nrhs + c];
            for (int p = i + 1; p < n; ++p) s = fma(-A[(long)p * n + i], Wout[(long)p * nrhs + c], s);
            Wout[(long)i * nrhs + c] = s / A[(long)i * n + i];
        }
    }
}

// ---- blocked Cholesky and solves for M + 1 > 126 ---------------------------
// (A0 + (ridge + jitter) I) W = B factored and solved by 64 x 64 blocks over
// many workgroups, one launch per step: for each diagonal block k the block
// is factored in LDS (k_bc_potrf), the panel below it solved against it
// (k_bc_trsm) and the trailing lower triangle updated (k_bc_syrk); then the
// forward and backward substitutions run block by block, a triangular solve
// of the block's rows (k_bc_fwd / k_bc_bwd) and a product that moves its
// contribution into the remaining rows (k_bc_fwd_update / k_bc_bwd_update).
// A non-positive pivot clears scal[kScalSolveOK]; every later kernel of the
// sequence then returns at once, and the host retries with a jitter.
constexpr int BC = 64;

__global__ __launch_bounds__(256) void k_bc_init(const double *__restrict__ A0,
                                                 double *__restrict__ A, int n, double ridge,
                                                 double jit, const double *__restrict__ B,
                                                 double *__restrict__ X, int nrhs,
                                                 double *__restrict__ scal) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long nn = (long)n * n;
    if (e < nn) A[e] = A0[e] + ((e / n == e % n) ? ridge + jit : 0.0);
    else if (e < nn + (long)n * nrhs) X[e - nn] = B[e - nn];
    if (e == 0) {
        scal[kScalSolveOK] = 1.0;
        scal[kScalJitter] = jit;
    }
}

// factor the kb x kb diagonal block at k0 in LDS (right-looking, column by column)
__global__ __launch_bounds__(256) void k_bc_potrf(double *__restrict__ A, int n, int k0,
                                                  double *__restrict__ scal) {
    __shared__ double L[BC][BC + 1];
    __shared__ int bad;
    if (scal[kScalSolveOK] != 1.0) return;
    const int tid = threadIdx.x;
    const int kb = min(BC, n - k0);
    for (int e = tid; e < kb * kb; e += 256) L[e / kb][e % kb] = A[(long)(k0 + e / kb) * n + k0 + e % kb];
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        if (tid == 0) {
            const double p = L[j][j];
            if (!(p > 0.0)) bad = 1;
            else L[j][j] = sqrt(p);
        }
        __syncthreads();
        if (bad) break;
        const double ljj = L[j][j];
        for (int i = j + 1 + tid; i < kb; i += 256) L[i][j] /= ljj;
        __syncthreads();
        const int m = kb - j - 1;
        for (int e = tid; e < m * m; e += 256) {
            const int r = j + 1 + e / m, c = j + 1 + e % m;
            if (c <= r) L[r][c] = fma(-L[r][j], L[c][j], L[r][c]);
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) scal[kScalSolveOK] = 0.0;
        return;
    }
    for (int e = tid; e < kb * kb; e += 256) {
        const int r = e / kb, c = e % kb;
        if (c <= r) A[(long)(k0 + r) * n + k0 + c] = L[r][c];
    }
}

// rows i >= k0 + kb of the panel: L21 = A21 L11^-T, one row per thread
__global__ __launch_bounds__(256) void k_bc_trsm(double *__restrict__ A, int n, int k0,
                                                 const double *__restrict__ scal) {
    __shared__ double L[BC][BC + 1];
    if (scal[kScalSolveOK] != 1.0) return;
    const int tid = threadIdx.x;
    const int kb = min(BC, n - k0), k1 = k0 + kb;
    for (int e = tid; e < kb * kb; e += 256) L[e / kb][e % kb] = A[(long)(k0 + e / kb) * n + k0 + e % kb];
    __syncthreads();
    const long i = (long)k1 + (long)blockIdx.x * 256 + tid;
    if (i >= n) return;
    double x[BC];
#pragma unroll
    for (int j = 0; j < BC; ++j) x[j] = (j < kb) ? A[i * n + k0 + j] : 0.0;
#pragma unroll
    for (int j = 0; j < BC; ++j) {
        if (j < kb) {
            double s = x[j];
#pragma unroll
            for (int p = 0; p < j; ++p) s = fma(-x[p], L[j][p], s);
            x[j] = s / L[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < BC; ++j)
        if (j < kb) A[i * n + k0 + j] = x[j];
}

// trailing update of the lower triangle: A[i][j] -= sum_p L[i][p] L[j][p]
// over the panel columns p, for k1 <= j <= i < n; 64 x 64 tiles (ti >= tj)
__global__ __launch_bounds__(256) void k_bc_syrk(double *__restrict__ A, int n, int k0,
                                                 const double *__restrict__ scal) {
    __shared__ double Li[BC][BC + 1], Lj[BC][BC + 1];
    if (blockIdx.y > blockIdx.x || scal[kScalSolveOK] != 1.0) return;
    const int kb = min(BC, n - k0), k1 = k0 + kb;
    const int i0 = k1 + blockIdx.x * BC, j0 = k1 + blockIdx.y * BC;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (int e = tid; e < BC * BC; e += 256) {
        const int r = e / BC, p = e % BC;
        Li[r][p] = (i0 + r < n && p < kb) ? A[(long)(i0 + r) * n + k0 + p] : 0.0;
        Lj[r][p] = (j0 + r < n && p < kb) ? A[(long)(j0 + r) * n + k0 + p] : 0.0;
    }
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int p = 0; p < kb; ++p) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = Li[ty + 16 * a][p];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = Lj[tx + 16 * b][p];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            if (i < n && j <= i) A[(long)i * n + j] -= acc[a][b];
        }
}

// forward substitution of block k: X[k0..k1) = L_kk^-1 X[k0..k1), one
// right-hand side per thread
__global__ __launch_bounds__(256) void k_bc_fwd(const double *__restrict__ A, int n, int k0,
                                                double *__restrict__ X, int nrhs,
                                                const double *__restrict__ scal) {
    __shared__ double L[BC][BC + 1];
    if (scal[kScalSolveOK] != 1.0) return;
    const int tid = threadIdx.x;
    const int kb = min(BC, n - k0);
    for (int e = tid; e < kb * kb; e += 256) L[e / kb][e % kb] = A[(long)(k0 + e / kb) * n + k0 + e % kb];
    __syncthreads();
    const int c = blockIdx.x * 256 + tid;
    if (c >= nrhs) return;
    double x[BC];
#pragma unroll
    for (int j = 0; j < BC; ++j) {
        if (j < kb) {
            double s = X[(long)(k0 + j) * nrhs + c];
#pragma unroll
            for (int p = 0; p < j; ++p) s = fma(-L[j][p], x[p], s);
            x[j] = s / L[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < BC; ++j)
        if (j < kb) X[(long)(k0 + j) * nrhs + c] = x[j];
}

// backward substitution of block k: X[k0..k1) = L_kk^-T X[k0..k1)
__global__ __launch_bounds__(256) void k_bc_bwd(const double *__restrict__ A, int n, int k0,
                                                double *__restrict__ X, int nrhs,
                                                const double *__restrict__ scal) {
    __shared__ double L[BC][BC + 1];
    if (scal[kScalSolveOK] != 1.0) return;
    const int tid = threadIdx.x;
    const int kb = min(BC, n - k0);
    for (int e = tid; e < kb * kb; e += 256) L[e / kb][e % kb] = A[(long)(k0 + e / kb) * n + k0 + e % kb];
    __syncthreads();
    const int c = blockIdx.x * 256 + tid;
    if (c >= nrhs) return;
    double x[BC];
#pragma unroll
    for (int jj = BC - 1; jj >= 0; --jj) {
        if (jj < kb) {
            double s = X[(long)(k0 + jj) * nrhs + c];
#pragma unroll
            for (int p = jj + 1; p < BC; ++p)
                if (p < kb) s = fma(-L[p][jj], x[p], s);
            x[jj] = s / L[jj][jj];
        }
    }
#pragma unroll
    for (int j = 0; j < BC; ++j)
        if (j < kb) X[(long)(k0 + j) * nrhs + c] = x[j];
}

// X[i] -= sum_p L[i][k0 + p] X[k0 + p] for the rows i >= k1 (forward) or
// X[i] -= sum_p L[k0 + p][i] X[k0 + p] for the rows i < k0 (backward), in
// 64 x 64 tiles of (rows, right-hand sides)
template <bool FWD>
__global__ __launch_bounds__(256) void k_bc_update(const double *__restrict__ A, int n, int k0,
                                                   double *__restrict__ X, int nrhs,
                                                   const double *__restrict__ scal) {
    __shared__ double Lt[BC][BC + 1], Xs[BC][BC + 1];
    if (scal[kScalSolveOK] != 1.0) return;
    const int kb = min(BC, n - k0), k1 = k0 + kb;
    const int r0 = (FWD ? k1 : 0) + blockIdx.x * BC, c0 = blockIdx.y * BC;
    const int rend = FWD ? n : k0;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (int e = tid; e < BC * BC; e += 256) {
        const int r = e / BC, p = e % BC;
        const int i = r0 + r;
        Lt[r][p] = (i < rend && p < kb)
                       ? (FWD ? A[(long)i * n + k0 + p] : A[(long)(k0 + p) * n + i]) : 0.0;
        Xs[r][p] = (r < kb && c0 + p < nrhs) ? X[(long)(k0 + r) * nrhs + c0 + p] : 0.0;
    }
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int p = 0; p < kb; ++p) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = Lt[ty + 16 * a][p];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = Xs[p][tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = r0 + ty + 16 * a, c = c0 + tx + 16 * b;
            if (i < rend && c < nrhs) X[(long)i * nrhs + c] -= acc[a][b];
        }
}

// tr(A0) into scal[kScalJitter] (the retries' jitter scale)
__global__ __launch_bounds__(64) void k_bc_trace(const double *__restrict__ A0, int n,
                                                 double *__restrict__ scal) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) s += A0[(long)i * n + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) scal[kScalJitter] = s;
}

// Tiled FP64 products of the M-step: 64 x 64 output tiles, 256 threads with
// 4 x 4 results each (rows ty + 16 a, columns tx + 16 b: conflict-free LDS
// reads), the inner dimension in steps of 16 staged through LDS.
constexpr int GT = 64, GK = 16;

// Split-K partials of [A | B] = FI^T [diag(G) FI | RT], an M1 x (M1 + D)
// product over the centres: workgroup (bx, by, c) computes one output tile
// over centres [c kc, min(K, (c + 1) kc)); part is nch x M1 x (M1 + D).
__device__ __forceinline__ void k_partAB_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ FI,
    const double *__restrict__ G, const double *__restrict__ RT, int K, int M1, int D, int kc,
    double *__restrict__ part) {
    __shared__ double As[GK][GT], Bs[GK][GT];
    const int NCOL = M1 + D;
    const int i0 = bx * GT, j0 = by * GT, c = bz;
    const int k0 = c * kc, k1 = (k0 + kc < K) ? k0 + kc : K;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    // the next step's operands are loaded into registers while this step's
    // are multiplied from LDS (four elements of each per thread)
    static_assert(GK * GT == 4 * 256, "four staged elements per thread");
    double ra[4], rb[4];
    auto gload = [&](int kb) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int kk = e / GT, col = e % GT, k = kb + kk;
            double av = 0.0, bv = 0.0;
            if (k < k1) {
                const int i = i0 + col, jj = j0 + col;
                if (i < M1) av = FI[(long)k * M1 + i];
                if (jj < M1) bv = G[k] * FI[(long)k * M1 + jj];
                else if (jj < NCOL) bv = RT[(long)k * D + (jj - M1)];
            }
            ra[u] = av;
            rb[u] = bv;
        }
    };
    if (k0 < k1) gload(k0);
    for (int kb = k0; kb < k1; kb += GK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            As[e / GT][e % GT] = ra[u];
            Bs[e / GT][e % GT] = rb[u];
        }
        __syncthreads();
        if (kb + GK < k1) gload(kb + GK);
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[kk][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
    const long E = (long)M1 * NCOL;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            if (i < M1 && j < NCOL) part[(long)c * E + (long)i * NCOL + j] = acc[a][b];
        }
}

__global__ __launch_bounds__(256) void k_partAB(const double *__restrict__ FI,
                                                const double *__restrict__ G,
                                                const double *__restrict__ RT, int K, int M1,
                                                int D, int kc, double *__restrict__ part) {
    k_partAB_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, FI, G, RT, K, M1, D, kc, part);
}

// A0 / B from the chunk partials, summed in chunk order.
__device__ __forceinline__ void k_sumAB_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ part,
    int nch, int M1, int D, double *__restrict__ A0, double *__restrict__ B) {
    const int NCOL = M1 + D;
    const long E = (long)M1 * NCOL;
    const long e = (long)bx * blockDim.x + threadIdx.x;
    if (e >= E) return;
    // eight loads in flight, then their sums in chunk order
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= nch; c += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(long)(c + u) * E + e];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < nch; ++c) s += part[(long)c * E + e];
    const int i = (int)(e / NCOL), j = (int)(e % NCOL);
    if (j < M1) A0[(long)i * M1 + j] = s;
    else B[(long)i * D + (j - M1)] = s;
}

__global__ void k_sumAB(const double *__restrict__ part, int nch, int M1, int D,
                        double *__restrict__ A0, double *__restrict__ B) {
    k_sumAB_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, part, nch, M1, D, A0, B);
}

constexpr int kLdsN = 126;   // largest M+1 whose factor fits in LDS
static_assert(kLdsN <= 128, "the Cholesky and solve kernels give each thread at most 4 rows");

// Broadcast lane l's value of x (l wave-uniform) to the whole wavefront.
__device__ __forceinline__ double lane_bcast(double x, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

constexpr int kPB = 8;        // panel width of the blocked Cholesky

// Cholesky of A0 + (ridge + jitter) I with the matrix resident in LDS (one
// workgroup, n <= 126); same pivots and retries as k_chol_solve.  Blocked,
// right-looking: wave 0 factors each 8-column panel in registers (lane r
// holds rows p0 + r and p0 + r + 64; the pivot row and the L entries of later
// panel columns are broadcast with readlane, so the panel needs no barrier),
// then all waves apply the panel to the trailing lower triangle with 4 x 4
// register tiles.  Two barriers per panel.  Writes the factor's lower
// triangle (the rest of L is left unspecified).
__device__ __forceinline__ void k_chol_lds_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ A0,
    double *__restrict__ L, int n, double ridge, double *__restrict__ scal) {
    __shared__ double As[kLdsN * kLdsN];
    __shared__ double Pt[kPB][128];                  // the factored panel, transposed
    __shared__ double tr;
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int tx = tid & 31, ty = tid >> 5;
    double jit = 0.0;
    int ok = 0;
    for (int attempt = 0; attempt < 9 && !ok; ++attempt) {
        if (attempt == 1) {
            block_trace(A0, n, &tr);
            jit = 1e-13 * tr / n;
        } else if (attempt > 1) {
            jit *= 10.0;
        }
        for (int e = tid; e < n * n; e += nt) {
            const int i = e / n, j = e % n;
            As[e] = A0[e] + ((i == j) ? ridge + jit : 0.0);
        }
        if (tid == 0) bad = 0;
        __syncthreads();
        for (int p0 = 0; p0 < n; p0 += kPB) {
            const int pn = min(kPB, n - p0);
            if (wv == 0) {
                const int r0 = p0 + lane, r1 = r0 + 64;
                double a0[kPB], a1[kPB];
#pragma unroll
                for (int c = 0; c < kPB; ++c) {
                    a0[c] = (c < pn && r0 < n) ? As[r0 * n + p0 + c] : 0.0;
                    a1[c] = (c < pn && r1 < n) ? As[r1 * n + p0 + c] : 0.0;
                }
                bool failed = false;
#pragma unroll
                for (int j = 0; j < kPB; ++j) {
                    if (j < pn && !failed) {
                        const double piv = lane_bcast(a0[j], j);     // row p0 + j is lane j's
                        if (!(piv > 0.0)) {
                            failed = true;                           // the same in every lane
                        } else {
                            const double sq = sqrt(piv), inv = 1.0 / sq;
                            if (lane == j) a0[j] = sq;
                            else if (lane > j) a0[j] *= inv;
                            a1[j] *= inv;
#pragma unroll
                            for (int c = j + 1; c < kPB; ++c) {
                                if (c < pn) {
                                    const double lcj = lane_bcast(a0[j], c);   // L[p0 + c][p0 + j]
                                    if (lane > j) a0[c] = fma(-a0[j], lcj, a0[c]);
                                    a1[c] = fma(-a1[j], lcj, a1[c]);
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < kPB; ++c)
                    if (c < pn) {
                        if (r0 < n) { As[r0 * n + p0 + c] = a0[c]; Pt[c][r0 - p0] = a0[c]; }
                        if (r1 < n) { As[r1 * n + p0 + c] = a1[c]; Pt[c][r1 - p0] = a1[c]; }
                    }
                if (failed && lane == 0) bad = 1;
            }
            __syncthreads();
            if (bad) break;                                  // read after the barrier: uniform
            // trailing update A[i][k] -= sum_j L[i][p0 + j] L[k][p0 + j] for
            // q0 <= k <= i: thread (ty, tx) owns rows q0 + ty + 32 a and columns
            // q0 + tx + 32 b, read from the transposed panel (consecutive lanes,
            // consecutive columns: no bank conflicts)
            const int q0 = p0 + pn;
            if (q0 + ty < n) {
                double acc[4][4];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
                const int oi = q0 - p0 + ty, ok0 = q0 - p0 + tx;   // panel-relative rows
                for (int j = 0; j < pn; ++j) {
                    double li[4], lk[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        li[a] = (oi + 32 * a < 128) ? Pt[j][oi + 32 * a] : 0.0;
                        lk[a] = (ok0 + 32 * a < 128) ? Pt[j][ok0 + 32 * a] : 0.0;
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) acc[a][b] = fma(li[a], lk[b], acc[a][b]);
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int i = q0 + ty + 32 * a, k = q0 + tx + 32 * b;
                        if (i < n && k <= i) As[i * n + k] -= acc[a][b];
                    }
            }
            __syncthreads();
        }
        ok = !bad;
        __syncthreads();                                     // before the next attempt resets bad
    }
    if (tid == 0) {
        scal[kScalJitter] = jit;
        scal[kScalSolveOK] = ok ? 1.0 : 0.0;
    }
    for (int e = tid; e < n * n; e += nt) L[e] = As[e];
}

__global__ __launch_bounds__(1024) void k_chol_lds(const double *__restrict__ A0,
                                                   double *__restrict__ L, int n, double ridge,
                                                   double *__restrict__ scal) {
    k_chol_lds_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, A0, L, n, ridge, scal);
}

// W[:, c0:c0+CPW NT/64] from L y = b, L^T x = y (n <= 126), NT threads: L in
// LDS (rows at an odd stride, so that a column read by a wave spreads over the
// banks); wave w solves the CPW right-hand sides c0 + CPW w + q on its own, lane
// l keeping rows l and l + 64 of each, and their 1 / L_pp, in registers.
// Step i: x_i and 1 / L_ii go from the lane that owns row i to the whole wave
// by readlane, with no barrier and no LDS access on the way, and every lane
// updates its rows after (forward) or before (backward) row i with the L
// elements it read from LDS during the step before.  The arithmetic is that
// of the plain substitution, x_i = x_i (1 / L_ii), then one FMA per later
// (earlier) row, in the same order for every row.
template <int NT, int CPW>
__device__ __forceinline__ void k_trsm_lds_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ L, int n,
    const double *__restrict__ B, int nrhs, double *__restrict__ W,
    const double *__restrict__ scal) {
    if (scal[kScalSolveOK] != 1.0) return;
    constexpr int LDL = kLdsN | 1;
    __shared__ double Ls[kLdsN * LDL];
    const int tid = threadIdx.x, nt = NT;
    const int l = tid & 63, w = tid >> 6;
    const int ld = n | 1;                          // row stride in LDS, at most LDL
    // the lower triangle: all of a thread's loads are in flight before its
    // first LDS store
    constexpr int FILL = (kLdsN * kLdsN + NT - 1) / NT;
    {
        double v[FILL];
#pragma unroll
        for (int f = 0; f < FILL; ++f) {
            const int e = tid + f * nt, i = e / n, k = e - i * n;
            v[f] = (e < n * n && k <= i) ? L[e] : 0.0;
        }
#pragma unroll
        for (int f = 0; f < FILL; ++f) {
            const int e = tid + f * nt, i = e / n, k = e - i * n;
            if (e < n * n && k <= i) Ls[i * ld + k] = v[f];
        }
    }
    const int c0 = bx * (CPW * NT / 64);
    double x[CPW][2], rdv[2];
    bool live[CPW];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int p = l + 64 * a;
        rdv[a] = (p < n) ? 1.0 / L[(long)p * n + p] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int col = c0 + CPW * w + q;
        live[q] = col < nrhs;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int p = l + 64 * a;
            x[q][a] = (live[q] && p < n) ? B[(long)p * nrhs + col] : 0.0;
        }
    }
    __syncthreads();
    // lane `src`'s v on every lane (src is the same for the whole wave)
    auto bcast = [](double v, int src) {
        const long long b = __builtin_bit_cast(long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
        return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
    };
    // x_i from its owner, finished
    auto pivot = [&](int i, double (&xi)[CPW]) {
        const int ai = i >> 6, li = i & 63;
        const double r = bcast(ai ? rdv[1] : rdv[0], li);
#pragma unroll
        for (int q = 0; q < CPW; ++q) xi[q] = bcast(ai ? x[q][1] : x[q][0], li) * r;
    };
    // the L elements of the next step (rows p < n only: the others lie
    // outside the LDS rows written)
    double lpn[2] = {0.0, 0.0};
    auto next_l = [&](int off, int pstride) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int p = l + 64 * a;
            if (p < n) lpn[a] = Ls[p * pstride + off];
        }
    };
    next_l(0, ld);                                 // column 0
    for (int i = 0; i < n; ++i) {                  // forward
        const double lp[2] = {lpn[0], lpn[1]};
        if (i + 1 < n) next_l(i + 1, ld);          // column i + 1
        double xi[CPW];
        pivot(i, xi);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int p = l + 64 * a;
            if (p == i) {
#pragma unroll
                for (int q = 0; q < CPW; ++q) x[q][a] = xi[q];
            } else if (p > i && p < n) {
#pragma unroll
                for (int q = 0; q < CPW; ++q) x[q][a] = fma(-lp[a], xi[q], x[q][a]);
            }
        }
    }
    next_l((n - 1) * ld, 1);                       // row n - 1
    for (int i = n - 1; i >= 0; --i) {             // backward with L^T
        const double lp[2] = {lpn[0], lpn[1]};
        if (i > 0) next_l((i - 1) * ld, 1);        // row i - 1
        double xi[CPW];
        pivot(i, xi);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int p = l + 64 * a;
            if (p == i) {
#pragma unroll
                for (int q = 0; q < CPW; ++q) x[q][a] = xi[q];
            } else if (p < i) {
#pragma unroll
                for (int q = 0; q < CPW; ++q) x[q][a] = fma(-lp[a], xi[q], x[q][a]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int col = c0 + CPW * w + q;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int p = l + 64 * a;
            if (live[q] && p < n) W[(long)p * nrhs + col] = x[q][a];
        }
    }
}

// one wave per SIMD and one right-hand side per wave: the substitution is
// issue-bound, so the right-hand sides go 4 to a workgroup over more CUs
constexpr int kTrsmNT = 256, kTrsmCPW = 1;
__global__ __launch_bounds__(kTrsmNT) void k_trsm_lds(const double *__restrict__ L, int n,
                                                      const double *__restrict__ B, int nrhs,
                                                      double *__restrict__ W,
                                                      const double *__restrict__ scal) {
    k_trsm_lds_body<kTrsmNT, kTrsmCPW>(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, L, n, B,
                                       nrhs, W, scal);
}

// Y = FI W into the padded FP64 layout (K x ld, columns >= D untouched):
// 64 x 64 output tiles over (centre, d), the M1 inner dimension in steps of 16.
__device__ __forceinline__ void k_phiw_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ FI,
    const double *__restrict__ W, int K, int M1, int D, int ld, double *__restrict__ Y) {
    __shared__ double As[GK][GT + 1], Bs[GK][GT];
    const int k0 = bx * GT, d0 = by * GT;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    // the next step's operands are loaded into registers while this step's
    // are multiplied from LDS (four elements of each per thread)
    static_assert(GK * GT == 4 * 256, "four staged elements per thread");
    double ra[4], rb[4];
    auto gload = [&](int ib) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int row = e / GK, ii = e % GK;          // FI rows read along i
            const int k = k0 + row, i = ib + ii;
            ra[u] = (k < K && i < M1) ? FI[(long)k * M1 + i] : 0.0;
            const int kk = e / GT, col = e % GT;          // W rows read along d
            const int i2 = ib + kk, d = d0 + col;
            rb[u] = (i2 < M1 && d < D) ? W[(long)i2 * D + d] : 0.0;
        }
    };
    gload(0);
    for (int ib = 0; ib < M1; ib += GK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            As[e % GK][e / GK] = ra[u];
            Bs[e / GT][e % GT] = rb[u];
        }
        __syncthreads();
        if (ib + GK < M1) gload(ib + GK);
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[kk][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = k0 + ty + 16 * a, d = d0 + tx + 16 * b;
            if (k < K && d < D) Y[(long)k * ld + d] = acc[a][b];
        }
}

__global__ __launch_bounds__(256) void k_phiw(const double *__restrict__ FI,
                                              const double *__restrict__ W, int K, int M1, int D,
                                              int ld, double *__restrict__ Y) {
    k_phiw_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, FI, W, K, M1, D, ld, Y);
}

// Per-workgroup partials of sum_k (G_k |y_k|^2 - 2 y_k . RT_k) over 64
// centres each (one wave per centre, 16 centres per wave); k_beta_final sums
// them in workgroup order into scal[kScalBetaSum].
constexpr int kBetaWG = 64;
__device__ __forceinline__ void k_beta_part_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ Y,
    int ld, const double *__restrict__ yy, const double *__restrict__ G,
    const double *__restrict__ RT, int K, int D, double *__restrict__ part) {
    __shared__ double s[4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double acc = 0.0;                              // lane 0: this wave's sum
    const int kw = bx * kBetaWG + w * (kBetaWG / 4);
    constexpr int DM = 3;
    if (D <= 64 * DM) {
        // the next centre's operands are loaded while this one's dot product
        // is reduced (same products and sums as the loop below)
        double yv[DM], rv[DM], gy[2];
        auto load = [&](int k) {
#pragma unroll
            for (int m = 0; m < DM; ++m) {
                const int d = l + 64 * m;
                const bool in = k < K && d < D;
                yv[m] = in ? Y[(long)k * ld + d] : 0.0;
                rv[m] = in ? RT[(long)k * D + d] : 0.0;
            }
            gy[0] = (k < K) ? G[k] : 0.0;
            gy[1] = (k < K) ? yy[k] : 0.0;
        };
        load(kw);
        for (int q = 0; q < kBetaWG / 4; ++q) {
            const int k = kw + q;
            if (k >= K) break;
            double y[DM], r[DM];
#pragma unroll
            for (int m = 0; m < DM; ++m) { y[m] = yv[m]; r[m] = rv[m]; }
            const double gk = gy[0], yk = gy[1];
            if (q + 1 < kBetaWG / 4) load(k + 1);
            double dot = 0.0;
#pragma unroll
            for (int m = 0; m < DM; ++m)
                if (l + 64 * m < D) dot = fma(y[m], r[m], dot);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
            acc += gk * yk - 2.0 * dot;
        }
    } else {
        for (int q = 0; q < kBetaWG / 4; ++q) {
            const int k = kw + q;
            if (k >= K) break;
            double dot = 0.0;
            for (int d = l; d < D; d += 64) dot = fma(Y[(long)k * ld + d], RT[(long)k * D + d], dot);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
            acc += G[k] * yy[k] - 2.0 * dot;
        }
    }
    if (l == 0) s[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[bx] = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ __launch_bounds__(256) void k_beta_part(const double *__restrict__ Y, int ld,
                                                   const double *__restrict__ yy,
                                                   const double *__restrict__ G,
                                                   const double *__restrict__ RT, int K, int D,
                                                   double *__restrict__ part) {
    k_beta_part_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, Y, ld, yy, G, RT, K, D, part);
}

__device__ __forceinline__ void k_beta_final_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ part,
    int n, double *__restrict__ scal) {
    __shared__ double s[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a += part[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) scal[kScalBetaSum] = s[0];
}

__global__ __launch_bounds__(256) void k_beta_final(const double *__restrict__ part, int n,
                                                    double *__restrict__ scal) {
    k_beta_final_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, part, n, scal);
}

// ---------------------------------------------------------------- batched
// The M-step of many instances per launch: tab holds one BInst per instance,
// list the table indices of the instances in this launch, and one grid
// dimension selects list[slot].  Each instance keeps its own grid extents;
// workgroups past them return before any barrier.  Every instance computes
// exactly what its own single-instance launch computes.

static_assert(GT == kBGT && kBetaWG == kBBetaWG, "flat-grid tile sizes");

// Flat-grid kernels: workgroup blockIdx.x belongs to instance map[.].x and is
// its workgroup map[.].y (grids sized by bk_count for each instance).
__global__ __launch_bounds__(256) void kb_copy_g(const BInst *__restrict__ tab,
                                                 const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    const int k = m.y * 256 + threadIdx.x;
    if (I.do_m && k < I.K) I.Gm[k] = I.G[k];
}

__global__ __launch_bounds__(256) void kb_partAB(const BInst *__restrict__ tab,
                                                 const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    const int nx = (I.M1 + GT - 1) / GT, ny = (I.M1 + I.D + GT - 1) / GT;
    const int kc = (I.K + I.nchAB - 1) / I.nchAB;
    k_partAB_body(m.y % nx, (m.y / nx) % ny, m.y / (nx * ny), 0, I.FI, I.G, I.RT, I.K, I.M1, I.D,
                  kc, I.partAB);
}

__global__ void kb_sumAB(const BInst *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    k_sumAB_body(m.y, 0, 0, 0, I.partAB, I.nchAB, I.M1, I.D, I.A0, I.Bm);
}

__global__ __launch_bounds__(1024) void kb_chol_lds(const BInst *__restrict__ tab,
                                                    const int *__restrict__ list) {
    const BInst &I = tab[list[blockIdx.x]];
    if (!I.do_m) return;
    k_chol_lds_body(0, 0, 0, 1, I.A0, I.A, I.M1, I.ridge, I.scal);
}

__global__ __launch_bounds__(1024) void kb_trsm_lds(const BInst *__restrict__ tab,
                                                    const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    k_trsm_lds_body<1024, 2>(m.y, 0, 0, 0, I.A, I.M1, I.Bm, I.D, I.W, I.scal);
}

__global__ __launch_bounds__(1024) void kb_chol_solve(const BInst *__restrict__ tab,
                                                      const int *__restrict__ list) {
    const BInst &I = tab[list[blockIdx.x]];
    if (!I.do_m) return;
    k_chol_solve_body(0, 0, 0, 1, I.A0, I.A, I.M1, I.ridge, I.Bm, I.D, I.W, I.scal);
}

__global__ __launch_bounds__(256) void kb_phiw(const BInst *__restrict__ tab,
                                               const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    const int nx = (I.K + GT - 1) / GT;
    k_phiw_body(m.y % nx, m.y / nx, 0, 0, I.FI, I.W, I.K, I.M1, I.D, I.ldt, I.Y64);
}

__global__ void kb_rownorm_y(const BInst *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    k_rownorm_body(m.y, 0, 0, 0, I.Y64, I.K, I.ldt, I.D, I.yy);
}

__global__ __launch_bounds__(256) void kb_beta_part(const BInst *__restrict__ tab,
                                                    const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    k_beta_part_body(m.y, 0, 0, 0, I.Y64, I.ldt, I.yy, I.G, I.RT, I.K, I.D, I.partAB);
}

__global__ __launch_bounds__(256) void kb_beta_final(const BInst *__restrict__ tab,
                                                     const int *__restrict__ list) {
    const BInst &I = tab[list[blockIdx.x]];
    if (!I.do_m) return;
    k_beta_final_body(0, 0, 0, 1, I.partAB, bk_count(I, BK_BETA_PART), I.scal);
}

}  // namespace

// dst[r * ld + d] = src[r * D + d] for d < D, 0 for D <= d < ld.
__global__ __launch_bounds__(256) void k_pad_rows(const double *__restrict__ src, long rows,
                                                  int D, long lds, int ld,
                                                  double *__restrict__ dst) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * ld) return;
    const long r = e / ld;
    const int d = (int)(e % ld);
    dst[e] = (d < D) ? src[r * lds + d] : 0.0;
}

int dg_pad_rows(dg_ctx *c, const double *src, long rows, int D, long lds, int ld, double *dst) {
    const long n = rows * ld;
    hipLaunchKernelGGL(k_pad_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, src,
                       rows, D, lds, ld, dst);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_rownorms(dg_ctx *c, const double *X, long rows, int ld, int D, double *out) {
    hipLaunchKernelGGL(k_rownorm, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, c->stream,
                       X, rows, ld, D, out);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_reduce_logs(dg_ctx *c) {
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(1024), 0, c->stream, c->logg, c->logh, c->N,
                       c->scal, (int)kScalSumLogG);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_count_zero, dim3(1), dim3(1024), 0, c->stream, c->pm, c->N, c->scal);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_sum_into(dg_ctx *c, const double *v, long n, int slot) {
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(1024), 0, c->stream, v, (const double *)nullptr, n,
                       c->scal, slot);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_centres_from_w(dg_ctx *c) {
    const int D = c->D, K = c->K;
    hipLaunchKernelGGL(k_phiw, dim3((unsigned)((K + GT - 1) / GT), (unsigned)((D + GT - 1) / GT)),
                       dim3(256), 0, c->stream, c->FI, c->W, K, c->M1, D, c->ldt, c->Y64);
    DG_HIP(c, hipGetLastError());
    return dg_rownorms(c, c->Y64, K, c->ldt, D, c->yy);
}

// (A0 + (ridge + jit) I) W = Bm for M + 1 > 126 by the blocked kernels;
// clears scal[kScalSolveOK] on a non-positive pivot
int blocked_chol_solve(dg_ctx *c, double ridge, double jit) {
    const int n = c->M1, D = c->D;
    const long tot = (long)n * n + (long)n * D;
    hipLaunchKernelGGL(k_bc_init, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream,
                       c->A0, c->A, n, ridge, jit, c->Bm, c->W, D, c->scal);
    const int nb = (n + BC - 1) / BC;
    for (int k = 0; k < nb; ++k) {
        const int k0 = k * BC, rest = n - k0 - BC;
        hipLaunchKernelGGL(k_bc_potrf, dim3(1), dim3(256), 0, c->stream, c->A, n, k0, c->scal);
        if (rest > 0) {
            const int tb = (rest + BC - 1) / BC;
            hipLaunchKernelGGL(k_bc_trsm, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0,
                               c->stream, c->A, n, k0, c->scal);
            hipLaunchKernelGGL(k_bc_syrk, dim3((unsigned)tb, (unsigned)tb), dim3(256), 0, c->stream,
                               c->A, n, k0, c->scal);
        }
    }
    const unsigned rb = (unsigned)((D + BC - 1) / BC);
    for (int k = 0; k < nb; ++k) {
        const int k0 = k * BC, rest = n - k0 - BC;
        hipLaunchKernelGGL(k_bc_fwd, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c->stream,
                           c->A, n, k0, c->W, D, c->scal);
        if (rest > 0)
            hipLaunchKernelGGL((k_bc_update<true>), dim3((unsigned)((rest + BC - 1) / BC), rb),
                               dim3(256), 0, c->stream, c->A, n, k0, c->W, D, c->scal);
    }
    for (int k = nb - 1; k >= 0; --k) {
        const int k0 = k * BC;
        hipLaunchKernelGGL(k_bc_bwd, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c->stream,
                           c->A, n, k0, c->W, D, c->scal);
        if (k0 > 0)
            hipLaunchKernelGGL((k_bc_update<false>), dim3((unsigned)((k0 + BC - 1) / BC), rb),
                               dim3(256), 0, c->stream, c->A, n, k0, c->W, D, c->scal);
    }
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_mstep_trace(dg_ctx *c) {
    hipLaunchKernelGGL(k_bc_trace, dim3(1), dim3(64), 0, c->stream, c->A0, c->M1, c->scal);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_mstep_run(dg_ctx *c, double jitter) {
    const int M1 = c->M1, D = c->D, K = c->K;
    const long E = (long)M1 * M1 + (long)M1 * D;
    const int nch = c->nchAB;
    const int kc = (K + nch - 1) / nch;
    hipLaunchKernelGGL(k_partAB,
                       dim3((unsigned)((M1 + GT - 1) / GT), (unsigned)((M1 + D + GT - 1) / GT),
                            (unsigned)nch),
                       dim3(256), 0, c->stream, c->FI, c->G, c->RT, K, M1, D, kc, c->partAB);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_sumAB, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, c->stream,
                       c->partAB, nch, M1, D, c->A0, c->Bm);
    DG_HIP(c, hipGetLastError());
    const double ridge = (c->cfg.lambda > 0.0) ? c->cfg.lambda / c->beta : 0.0;
    if (M1 <= kLdsN) {
        hipLaunchKernelGGL(k_chol_lds, dim3(1), dim3(1024), 0, c->stream, c->A0, c->A, M1, ridge,
                           c->scal);
        DG_HIP(c, hipGetLastError());
        constexpr int cols = kTrsmCPW * kTrsmNT / 64;     // right-hand sides per workgroup
        hipLaunchKernelGGL(k_trsm_lds, dim3((unsigned)((D + cols - 1) / cols)),
                           dim3(kTrsmNT), 0, c->stream,
                           c->A, M1, c->Bm, D, c->W, c->scal);
    } else {
        int rc = blocked_chol_solve(c, ridge, jitter);
        if (rc) return rc;
    }
    DG_HIP(c, hipGetLastError());
    int rc = dg_centres_from_w(c);
    if (rc) return rc;
    // beta partials reuse the (now consumed) A/B partial buffer
    const int nb = (K + kBetaWG - 1) / kBetaWG;
    hipLaunchKernelGGL(k_beta_part, dim3((unsigned)nb), dim3(256), 0, c->stream, c->Y64, c->ldt,
                       c->yy, c->G, c->RT, K, D, c->partAB);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_beta_final, dim3(1), dim3(256), 0, c->stream, c->partAB, nb, c->scal);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

// M-step of the instances in list (n of them; nlds with M + 1 <= 126 first in
// list, the others after), on one stream; gr holds the flat grids over that
// list (the LDS triangular solves over its first nlds entries).  Each
// instance's steps are those of dg_mstep_run after the copy of G into Gm.
int dg_batch_mstep(hipStream_t st, const BInst *tab, const int *list, int n, int nlds,
                   const BGrids &gr) {
    if (n <= 0) return DG_OK;
    auto flat = [&](auto kernel, int kind, int threads) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)gr.wg[kind]), dim3(threads), 0, st, tab,
                           gr.map[kind]);
    };
    flat(kb_copy_g, BK_COPY_G, 256);
    flat(kb_partAB, BK_PART_AB, 256);
    flat(kb_sumAB, BK_SUM_AB, 256);
    if (nlds > 0) {
        hipLaunchKernelGGL(kb_chol_lds, dim3((unsigned)nlds), dim3(1024), 0, st, tab, list);
        flat(kb_trsm_lds, BK_TRSM_LDS, 1024);
    }
    if (n > nlds)
        hipLaunchKernelGGL(kb_chol_solve, dim3((unsigned)(n - nlds)), dim3(1024), 0, st, tab,
                           list + nlds);
    flat(kb_phiw, BK_PHIW, 256);
    flat(kb_rownorm_y, BK_ROWNORM, 256);
    flat(kb_beta_part, BK_BETA_PART, 256);
    hipLaunchKernelGGL(kb_beta_final, dim3((unsigned)n), dim3(256), 0, st, tab, list);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? DG_OK : DG_EHIP;
}
