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
__global__ void k_rownorm(const double *__restrict__ X, long rows, int ld, int D,
                          double *__restrict__ out) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
        const double v = X[r * ld + d];
        s = fma(v, v, s);
    }
    out[r] = s;
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

// Cholesky solve of (A0 + (ridge + jitter) I) W = B in one workgroup, the
// factor built in place in A (right-looking; each entry receives its
// updates in the same column order as a left-looking loop).  On a
// non-positive pivot the factorisation restarts with jitter 1e-13 tr/n,
// growing x10 per retry, at most 8 retries.  scal[kScalSolveOK] = 1 on
// success, scal[kScalJitter] = the jitter used.
__global__ __launch_bounds__(1024) void k_chol_solve(const double *__restrict__ A0,
                                                     double *__restrict__ A, int n,
                                                     double ridge, const double *__restrict__ B,
                                                     int nrhs, double *__restrict__ Wout,
                                                     double *__restrict__ scal) {
    __shared__ int fail;
    __shared__ double tr;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += A0[(long)i * n + i];
        tr = s;
    }
    __syncthreads();
    double jit = 0.0;
    int ok = 0;
    for (int attempt = 0; attempt < 9 && !ok; ++attempt) {
        if (attempt == 1) jit = 1e-13 * tr / n;
        else if (attempt > 1) jit *= 10.0;
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

// Split-K partials of A = FI^T diag(G) FI and B = FI^T RT: chunk c covers
// centres [c*kc, min(K, (c+1)*kc)); entry e < M1*M1 is A[e / M1][e % M1],
// the rest B[(e - M1*M1) / D][(e - M1*M1) % D].  part is NCH x E.
__global__ void k_partAB(const double *__restrict__ FI, const double *__restrict__ G,
                         const double *__restrict__ RT, int K, int M1, int D, int kc,
                         double *__restrict__ part) {
    const long EA = (long)M1 * M1, E = EA + (long)M1 * D;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (e >= E) return;
    const int k0 = c * kc;
    const int k1 = (k0 + kc < K) ? k0 + kc : K;
    double s = 0.0;
    if (e < EA) {
        const int i = (int)(e / M1), j = (int)(e % M1);
        for (int k = k0; k < k1; ++k) {
            const double g = FI[(long)k * M1 + j] * G[k];
            s = fma(FI[(long)k * M1 + i], g, s);
        }
    } else {
        const long e2 = e - EA;
        const int i = (int)(e2 / D), d = (int)(e2 % D);
        for (int k = k0; k < k1; ++k) s = fma(FI[(long)k * M1 + i], RT[(long)k * D + d], s);
    }
    part[(long)c * E + e] = s;
}

// A0 / B from the chunk partials, summed in chunk order.
__global__ void k_sumAB(const double *__restrict__ part, int nch, int M1, int D,
                        double *__restrict__ A0, double *__restrict__ B) {
    const long EA = (long)M1 * M1, E = EA + (long)M1 * D;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[(long)c * E + e];
    if (e < EA) A0[e] = s;
    else B[e - EA] = s;
}

constexpr int kLdsN = 126;   // largest M+1 whose factor fits in LDS

// Cholesky of A0 + (ridge + jitter) I with the matrix resident in LDS (one
// workgroup, n <= 126); same pivots, retries and update order as
// k_chol_solve.  Writes the factor's lower triangle to L.
__global__ __launch_bounds__(1024) void k_chol_lds(const double *__restrict__ A0,
                                                   double *__restrict__ L, int n, double ridge,
                                                   double *__restrict__ scal) {
    __shared__ double As[kLdsN * kLdsN];
    __shared__ int fail;
    __shared__ double tr;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += A0[(long)i * n + i];
        tr = s;
    }
    __syncthreads();
    double jit = 0.0;
    int ok = 0;
    for (int attempt = 0; attempt < 9 && !ok; ++attempt) {
        if (attempt == 1) jit = 1e-13 * tr / n;
        else if (attempt > 1) jit *= 10.0;
        for (int e = tid; e < n * n; e += nt) {
            const int i = e / n, j = e % n;
            As[e] = A0[e] + ((i == j) ? ridge + jit : 0.0);
        }
        if (tid == 0) fail = 0;
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            if (tid == 0) {
                const double p = As[j * n + j];
                if (!(p > 0.0)) fail = 1;
                else As[j * n + j] = sqrt(p);
            }
            __syncthreads();
            if (fail) break;
            const double ljj = As[j * n + j];
            for (int i = j + 1 + tid; i < n; i += nt) As[i * n + j] /= ljj;
            __syncthreads();
            const int m = n - j - 1;
            const int pairs = m * (m + 1) / 2;
            for (int e = tid; e < pairs; e += nt) {
                int r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
                while ((r + 1) * (r + 2) / 2 <= e) ++r;
                while (r * (r + 1) / 2 > e) --r;
                const int cc = e - r * (r + 1) / 2;
                const int i = j + 1 + r, k = j + 1 + cc;
                As[i * n + k] = fma(-As[i * n + j], As[k * n + j], As[i * n + k]);
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
    for (int e = tid; e < n * n; e += nt) L[e] = As[e];
}

// W[:, c0:c0+32] from L y = b, L^T x = y with L and the 32 right-hand sides
// in LDS (n <= 126).  Forward and backward substitution are right-looking
// over the 32 columns in parallel.
__global__ __launch_bounds__(256) void k_trsm_lds(const double *__restrict__ L, int n,
                                                  const double *__restrict__ B, int nrhs,
                                                  double *__restrict__ W,
                                                  const double *__restrict__ scal) {
    if (scal[kScalSolveOK] != 1.0) return;
    __shared__ double Ls[kLdsN * kLdsN];
    __shared__ double Xs[kLdsN * 32];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int c0 = blockIdx.x * 32;
    const int nc = (nrhs - c0 < 32) ? nrhs - c0 : 32;
    for (int e = tid; e < n * n; e += nt) Ls[e] = L[e];
    for (int e = tid; e < n * 32; e += nt) {
        const int i = e / 32, cc = e % 32;
        Xs[e] = (cc < nc) ? B[(long)i * nrhs + c0 + cc] : 0.0;
    }
    __syncthreads();
    for (int i = 0; i < n; ++i) {                  // forward
        if (tid < 32) Xs[i * 32 + tid] /= Ls[i * n + i];
        __syncthreads();
        for (int e = tid; e < (n - i - 1) * 32; e += nt) {
            const int p = i + 1 + e / 32, cc = e % 32;
            Xs[p * 32 + cc] = fma(-Ls[p * n + i], Xs[i * 32 + cc], Xs[p * 32 + cc]);
        }
        __syncthreads();
    }
    for (int i = n - 1; i >= 0; --i) {             // backward with L^T
        if (tid < 32) Xs[i * 32 + tid] /= Ls[i * n + i];
        __syncthreads();
        for (int e = tid; e < i * 32; e += nt) {
            const int p = e / 32, cc = e % 32;
            Xs[p * 32 + cc] = fma(-Ls[i * n + p], Xs[i * 32 + cc], Xs[p * 32 + cc]);
        }
        __syncthreads();
    }
    for (int e = tid; e < n * 32; e += nt) {
        const int i = e / 32, cc = e % 32;
        if (cc < nc) W[(long)i * nrhs + c0 + cc] = Xs[e];
    }
}

// Y[k][d] = sum_i FI[k][i] W[i][d] into the padded FP64 layout.
__global__ void k_phiw(const double *__restrict__ FI, const double *__restrict__ W, int K,
                       int M1, int D, int ld, double *__restrict__ Y) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (d >= D) return;
    double s = 0.0;
    for (int i = 0; i < M1; ++i) s = fma(FI[(long)k * M1 + i], W[(long)i * D + d], s);
    Y[(long)k * ld + d] = s;
}

// Per-workgroup partials of sum_k (G_k |y_k|^2 - 2 y_k . RT_k) over 256
// centres each; k_beta_final sums them in workgroup order into
// scal[kScalBetaSum].
__global__ __launch_bounds__(256) void k_beta_part(const double *__restrict__ Y, int ld,
                                                   const double *__restrict__ yy,
                                                   const double *__restrict__ G,
                                                   const double *__restrict__ RT, int K, int D,
                                                   double *__restrict__ part) {
    __shared__ double s[256];
    const int k = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (k < K) {
        double dot = 0.0;
        for (int d = 0; d < D; ++d) dot = fma(Y[(long)k * ld + d], RT[(long)k * D + d], dot);
        acc = G[k] * yy[k] - 2.0 * dot;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_beta_final(const double *__restrict__ part, int n, double *__restrict__ scal) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        scal[kScalBetaSum] = s;
    }
}

}  // namespace

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

int dg_mstep_run(dg_ctx *c) {
    const int M1 = c->M1, D = c->D, K = c->K;
    const long E = (long)M1 * M1 + (long)M1 * D;
    const int nch = c->nchAB;
    const int kc = (K + nch - 1) / nch;
    hipLaunchKernelGGL(k_partAB, dim3((unsigned)((E + 255) / 256), (unsigned)nch), dim3(256), 0,
                       c->stream, c->FI, c->G, c->RT, K, M1, D, kc, c->partAB);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_sumAB, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, c->stream,
                       c->partAB, nch, M1, D, c->A0, c->Bm);
    DG_HIP(c, hipGetLastError());
    const double ridge = (c->cfg.lambda > 0.0) ? c->cfg.lambda / c->beta : 0.0;
    if (M1 <= kLdsN) {
        hipLaunchKernelGGL(k_chol_lds, dim3(1), dim3(1024), 0, c->stream, c->A0, c->A, M1, ridge,
                           c->scal);
        DG_HIP(c, hipGetLastError());
        hipLaunchKernelGGL(k_trsm_lds, dim3((unsigned)((D + 31) / 32)), dim3(256), 0, c->stream,
                           c->A, M1, c->Bm, D, c->W, c->scal);
    } else {
        hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(1024), 0, c->stream, c->A0, c->A, M1,
                           ridge, c->Bm, D, c->W, c->scal);
    }
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_phiw, dim3((unsigned)((D + 63) / 64), (unsigned)K), dim3(64), 0,
                       c->stream, c->FI, c->W, K, M1, D, c->ldt, c->Y64);
    DG_HIP(c, hipGetLastError());
    int rc = dg_rownorms(c, c->Y64, K, c->ldt, D, c->yy);
    if (rc) return rc;
    // beta partials reuse the (now consumed) A/B partial buffer
    const int nb = (K + 255) / 256;
    hipLaunchKernelGGL(k_beta_part, dim3((unsigned)nb), dim3(256), 0, c->stream, c->Y64, c->ldt,
                       c->yy, c->G, c->RT, K, D, c->partAB);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_beta_final, dim3(1), dim3(64), 0, c->stream, c->partAB, nb, c->scal);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}
