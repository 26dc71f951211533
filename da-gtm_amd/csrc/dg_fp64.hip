// FP64 E-step of DA-GTM on gfx950: two streaming passes over the K x N
// distance matrix, which is never stored.
//
// Pass 1 (k_e1_f64): one workgroup per 64-point tile walks all centres in
// 64-centre tiles.  Each thread keeps an online log-sum-exp state for 4
// points over its 4-centre slice of every tile; the 16 slices are merged in
// LDS at the end.  Output per point: x_max, log S2 (pass-2 normaliser),
// log g and log h.
//
// Pass 2 (k_e2_f64): one workgroup per (64-centre tile, N chunk) recomputes
// the same distances, forms R = exp((x - x_max)/temp - log S2) in LDS and
// accumulates G = sum_n R and RT = sum_n R t_n in registers, writing one
// partial per chunk; k_reduce_parts sums the chunks in a fixed order.
//
// The math restates mpicomm.c:316-414 (resp) and mpicomm.c:463-500 (qual):
// x = d * (-beta/2), E = exp(x), P = E^(1/temp), R = P / colsum(P), with
// exp underflow (x < -745.13) giving exactly 0 and an all-zero column
// falling back to R = 1/K (h = K; g = K in EM mode only).
#include <cmath>

#include "dg_internal.h"

namespace {

constexpr int NT = 256;   // threads per workgroup (4 waves of 64)
constexpr int TP = 64;    // points per tile
constexpr int TK = 64;    // centres per tile
constexpr int DC = 16;    // d-chunk of the cross-term loop

// S[dd][r] = X[(r0 + r) * ld + d0 + dd], zero for rows >= rows.
__device__ __forceinline__ void stage_chunk(double (*S)[TP + 1], const double *__restrict__ X,
                                            long rows, long r0, int ld, int d0, int tid) {
#pragma unroll
    for (int e = tid; e < TP * DC; e += NT) {
        const int r = e / DC, dd = e % DC;
        const long row = r0 + r;
        S[dd][r] = (row < rows) ? X[row * ld + d0 + dd] : 0.0;
    }
}

// acc[i][j] += sum over the chunk of Ts[.][tx + 16 i] * Ys[.][ty + 16 j]
__device__ __forceinline__ void cross_chunk(double acc[4][4], const double (*Ts)[TP + 1],
                                            const double (*Ys)[TK + 1], int tx, int ty) {
#pragma unroll
    for (int dd = 0; dd < DC; ++dd) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Ts[dd][tx + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Ys[dd][ty + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
}

// Exponent x = d * (-beta/2) with d = (-2 s + |y|^2) + |t|^2 in the
// reference's dgemm + dger order (comm.c:590-619); -inf when exp underflows.
__device__ __forceinline__ double exponent(double s, double yk, double tn, double nbh, bool valid) {
    const double d = fma(-2.0, s, yk) + tn;
    const double x = d * nbh;
    return (valid && x > kExpUnderflowX) ? x : -INFINITY;
}

// ---------------------------------------------------------------- pass 1
__global__ __launch_bounds__(NT) void k_e1_f64(
    const double *__restrict__ T, const double *__restrict__ tt, long N,
    const double *__restrict__ Y, const double *__restrict__ yy, int K, int ld,
    double nbh, double it, int two, int isDA, double logK,
    double *__restrict__ pm, double *__restrict__ pl,
    double *__restrict__ logg, double *__restrict__ logh) {
    __shared__ double Ts[DC][TP + 1];
    __shared__ double Ys[DC][TK + 1];
    __shared__ double red[3][16][TP];

    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const long n0 = (long)blockIdx.x * TP;

    double tn[4], xm[4], s1[4], s2[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long n = n0 + tx + 16 * i;
        tn[i] = (n < N) ? tt[n] : 0.0;
        xm[i] = -INFINITY; s1[i] = 0.0; s2[i] = 0.0;
    }

    for (int k0 = 0; k0 < K; k0 += TK) {
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int d0 = 0; d0 < ld; d0 += DC) {
            __syncthreads();
            stage_chunk(Ts, T, N, n0, ld, d0, tid);
            stage_chunk(Ys, Y, K, k0, ld, d0, tid);
            __syncthreads();
            cross_chunk(acc, Ts, Ys, tx, ty);
        }
        double yk[4]; bool kv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + ty + 16 * j;
            kv[j] = k < K;
            yk[j] = kv[j] ? yy[k] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double x[4], t = -INFINITY;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                x[j] = exponent(acc[i][j], yk[j], tn[i], nbh, kv[j]);
                t = fmax(t, x[j]);
            }
            if (t > xm[i]) {             // new running max: rescale the sums
                const double r = xm[i] - t;
                s1[i] *= exp(r);
                if (two) s2[i] *= exp(r * it);
                xm[i] = t;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (x[j] != -INFINITY) {
                    const double r = x[j] - xm[i];
                    s1[i] += exp(r);
                    if (two) s2[i] += exp(r * it);
                }
            }
        }
    }

#pragma unroll
    for (int i = 0; i < 4; ++i) {
        red[0][ty][tx + 16 * i] = xm[i];
        red[1][ty][tx + 16 * i] = s1[i];
        red[2][ty][tx + 16 * i] = s2[i];
    }
    __syncthreads();
    if (tid < TP) {
        const int p = tid;
        const long n = n0 + p;
        double M = -INFINITY;
        for (int t = 0; t < 16; ++t) M = fmax(M, red[0][t][p]);
        double S1 = 0.0, S2 = 0.0;
        if (M != -INFINITY) {
            for (int t = 0; t < 16; ++t) {
                const double m = red[0][t][p];
                if (m == -INFINITY) continue;
                const double r = m - M;
                S1 += red[1][t][p] * exp(r);
                if (two) S2 += red[2][t][p] * exp(r * it);
            }
        }
        if (n < N) {
            if (M == -INFINITY) {          // zero column (mpicomm.c:382-390)
                pm[n] = -INFINITY;
                pl[n] = 0.0;
                logh[n] = logK;
                logg[n] = isDA ? -INFINITY : logK;
            } else {
                const double lS1 = log(S1);
                const double lS2 = two ? log(S2) : lS1;
                pm[n] = M;
                pl[n] = lS2;
                logg[n] = M + lS1;
                logh[n] = M * it + lS2;
            }
        }
    }
}

// ---------------------------------------------------------------- pass 2
template <int DJ>
struct E2Smem {
    double Tb[TP][16 * DJ + 1];
    double Rs[TK][TP + 1];
    double Ys[DC][TK + 1];
    double gr[16][TK];
};

template <int DJ>
__global__ __launch_bounds__(NT) void k_e2_f64(
    const double *__restrict__ T, const double *__restrict__ tt, long N,
    const double *__restrict__ Y, const double *__restrict__ yy, int K, int D,
    double nbh, double it, double invK,
    const double *__restrict__ pm, const double *__restrict__ pl, long chunk,
    double *__restrict__ partRT, double *__restrict__ partG) {
    constexpr int ld = 16 * DJ;
    __shared__ E2Smem<DJ> sm;

    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int kq = tid >> 4, dq = tid & 15;          // RT-stage mapping
    const int k0 = blockIdx.x * TK;
    const long nb = (long)blockIdx.y * chunk;
    const long ne = (nb + chunk < N) ? nb + chunk : N;

    double yk[4]; bool kv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = k0 + ty + 16 * j;
        kv[j] = k < K;
        yk[j] = kv[j] ? yy[k] : 0.0;
    }
    double rt[4][DJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < DJ; ++j) rt[i][j] = 0.0;
    double gacc[4] = {0.0, 0.0, 0.0, 0.0};

    for (long n0 = nb; n0 < ne; n0 += TP) {
        __syncthreads();
        for (int e = tid; e < TP * ld; e += NT) {
            const int r = e / ld, d = e % ld;
            const long n = n0 + r;
            sm.Tb[r][d] = (n < ne) ? T[n * ld + d] : 0.0;
        }
        double tn[4], m[4], l[4]; bool nv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long n = n0 + tx + 16 * i;
            nv[i] = n < ne;
            tn[i] = nv[i] ? tt[n] : 0.0;
            m[i] = nv[i] ? pm[n] : 0.0;
            l[i] = nv[i] ? pl[n] : 0.0;
        }
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int d0 = 0; d0 < ld; d0 += DC) {
            __syncthreads();
            stage_chunk(sm.Ys, Y, K, k0, ld, d0, tid);
            __syncthreads();
#pragma unroll
            for (int dd = 0; dd < DC; ++dd) {
                double a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = sm.Tb[tx + 16 * i][d0 + dd];
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = sm.Ys[dd][ty + 16 * j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double r = 0.0;
                if (nv[i]) {
                    if (m[i] == -INFINITY) {
                        r = kv[j] ? invK : 0.0;
                    } else {
                        const double x = exponent(acc[i][j], yk[j], tn[i], nbh, kv[j]);
                        r = (x != -INFINITY) ? exp((x - m[i]) * it - l[i]) : 0.0;
                    }
                }
                gacc[j] += r;
                sm.Rs[ty + 16 * j][tx + 16 * i] = r;
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int n = 0; n < TP; ++n) {
            double rv[4], tv[DJ];
#pragma unroll
            for (int i = 0; i < 4; ++i) rv[i] = sm.Rs[kq + 16 * i][n];
#pragma unroll
            for (int j = 0; j < DJ; ++j) tv[j] = sm.Tb[n][dq + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < DJ; ++j) rt[i][j] = fma(rv[i], tv[j], rt[i][j]);
        }
    }

    // partial RT of this chunk
    const long c = blockIdx.y;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = k0 + kq + 16 * i;
        if (k >= K) continue;
#pragma unroll
        for (int j = 0; j < DJ; ++j) {
            const int d = dq + 16 * j;
            if (d < D) partRT[(c * K + k) * D + d] = rt[i][j];
        }
    }
    // partial G: reduce gacc over the 16 threads that share ty
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) sm.gr[tx][ty + 16 * j] = gacc[j];
    __syncthreads();
    if (tid < TK) {
        double s = 0.0;
        for (int t = 0; t < 16; ++t) s += sm.gr[t][tid];
        const int k = k0 + tid;
        if (k < K) partG[c * K + k] = s;
    }
}

__global__ void k_reduce_parts(const double *__restrict__ partRT, const double *__restrict__ partG,
                               int NC, int K, int D, double *__restrict__ RT,
                               double *__restrict__ G) {
    const long KD = (long)K * D;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < KD) {
        double s = 0.0;
        for (int c = 0; c < NC; ++c) s += partRT[c * KD + i];
        RT[i] = s;
    }
    if (i < K) {
        double s = 0.0;
        for (int c = 0; c < NC; ++c) s += partG[(long)c * K + i];
        G[i] = s;
    }
}

// ---------------------------------------------------------------- debug / MAP
// One thread per (k, n): R_kn from the pass-1 normaliser.
__global__ void k_resp_dense(const double *__restrict__ T, const double *__restrict__ tt, long N,
                             const double *__restrict__ Y, const double *__restrict__ yy, int K,
                             int ld, double nbh, double it, double invK,
                             const double *__restrict__ pm, const double *__restrict__ pl,
                             double *__restrict__ R, long ldR) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n >= N) return;
    double r;
    if (pm[n] == -INFINITY) {
        r = invK;
    } else {
        double s = 0.0;
        for (int d = 0; d < ld; ++d) s = fma(T[n * ld + d], Y[(long)k * ld + d], s);
        const double x = exponent(s, yy[k], tt[n], nbh, true);
        r = (x != -INFINITY) ? exp((x - pm[n]) * it - pl[n]) : 0.0;
    }
    R[(long)k * ldR + n] = r;
}

// MAP: M[n][l] = sum_k R_kn X[k][l], one thread per point (L <= 3).
__global__ void k_map_f64(const double *__restrict__ T, const double *__restrict__ tt, long N,
                          const double *__restrict__ Y, const double *__restrict__ yy, int K,
                          int ld, double nbh, double it, double invK,
                          const double *__restrict__ pm, const double *__restrict__ pl,
                          const double *__restrict__ X, int L, double *__restrict__ M) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double acc[3] = {0.0, 0.0, 0.0};
    const bool zc = pm[n] == -INFINITY;
    for (int k = 0; k < K; ++k) {
        double r;
        if (zc) {
            r = invK;
        } else {
            double s = 0.0;
            for (int d = 0; d < ld; ++d) s = fma(T[n * ld + d], Y[(long)k * ld + d], s);
            const double x = exponent(s, yy[k], tt[n], nbh, true);
            r = (x != -INFINITY) ? exp((x - pm[n]) * it - pl[n]) : 0.0;
        }
        for (int l = 0; l < L; ++l) acc[l] = fma(r, X[(long)k * L + l], acc[l]);
    }
    for (int l = 0; l < L; ++l) M[n * L + l] = acc[l];
}

template <int DJ>
int launch_e2(dg_ctx *c, double nbh, double it, long chunk, dim3 grid) {
    hipLaunchKernelGGL(k_e2_f64<DJ>, grid, dim3(NT), 0, c->stream, c->T64, c->tt, c->N, c->Y64,
                       c->yy, c->K, c->D, nbh, it, 1.0 / c->K, c->pm, c->pl, chunk, c->partRT,
                       c->partG);
    return DG_OK;
}

// ---------------------------------------------------------------- critical temperature
// Responsibility-weighted second moments of the data centred on one centre,
// S1 = sum_n r_kn x x^T and S2 = sum_n r_kn^2 x x^T with x = t_n - y_k (the
// 2 S1 and -S2 of mpicomm.c:580-597).  One workgroup per (64 x 64 tile of the
// lower triangle, centre k0 + blockIdx.y, N chunk); R_kn is recomputed with
// k_resp_dense's formula.  Each thread owns a 4 x 4 micro-tile of both sums,
// accumulated over the chunk's points in order.  The chunk's partial goes to
// part[((chunk * KB + kb) * 2 + s) * D * D + a * D + b] for the tile only.
constexpr int CM = 64;   // moment tile edge

__global__ __launch_bounds__(NT) void k_ctemp_moments(
    const double *__restrict__ T, const double *__restrict__ tt, long N,
    const double *__restrict__ Y, const double *__restrict__ yy, int k0, int D, int ld,
    double nbh, double it, double invK, const double *__restrict__ pm,
    const double *__restrict__ pl, long chunk, double *__restrict__ part) {
    __shared__ double Xa[CM][CM + 1];
    __shared__ double Xb[CM][CM + 1];
    __shared__ double Ys[192];
    __shared__ double rs[CM];
    const int tid = threadIdx.x;
    int p = blockIdx.x, ti = 0;            // lower-triangle pair p -> (ti >= tj)
    while (p > ti) { p -= ti + 1; ++ti; }
    const int tj = p;
    const int kb = blockIdx.y, k = k0 + kb;
    const long n0 = (long)blockIdx.z * chunk;
    const long n1 = min(N, n0 + chunk);
    const int a0 = ti * CM, b0 = tj * CM;
    const int tx = tid & 15, ty = tid >> 4;
    for (int d = tid; d < ld; d += NT) Ys[d] = Y[(long)k * ld + d];
    double s1[4][4], s2[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) s1[i][j] = s2[i][j] = 0.0;
    __syncthreads();
    for (long base = n0; base < n1; base += CM) {
        if (tid < CM) {
            const long n = base + tid;
            double r = 0.0;
            if (n < n1) {
                if (pm[n] == -INFINITY) {
                    r = invK;
                } else {
                    double s = 0.0;
                    for (int d = 0; d < ld; ++d) s = fma(T[n * ld + d], Ys[d], s);
                    const double x = exponent(s, yy[k], tt[n], nbh, true);
                    r = (x != -INFINITY) ? exp((x - pm[n]) * it - pl[n]) : 0.0;
                }
            }
            rs[tid] = r;
        }
        for (int e = tid; e < CM * CM; e += NT) {
            const int q = e / CM, dd = e % CM;
            const long n = base + q;
            const int a = a0 + dd, b = b0 + dd;
            Xa[q][dd] = (n < n1 && a < D) ? T[n * ld + a] - Ys[a] : 0.0;
            Xb[q][dd] = (n < n1 && b < D) ? T[n * ld + b] - Ys[b] : 0.0;
        }
        __syncthreads();
        const int np = (int)min((long)CM, n1 - base);
        for (int q = 0; q < np; ++q) {
            const double r = rs[q];
            double ra[4], xb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) ra[i] = Xa[q][tx + 16 * i] * r;
#pragma unroll
            for (int j = 0; j < 4; ++j) xb[j] = Xb[q][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double rra = ra[i] * r;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    s1[i][j] = fma(ra[i], xb[j], s1[i][j]);
                    s2[i][j] = fma(rra, xb[j], s2[i][j]);
                }
            }
        }
        __syncthreads();
    }
    const long DD = (long)D * D;
    double *o1 = part + ((long)blockIdx.z * gridDim.y + kb) * 2 * DD;
    double *o2 = o1 + DD;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int a = a0 + tx + 16 * i, b = b0 + ty + 16 * j;
            if (a < D && b < D) {
                o1[(long)a * D + b] = s1[i][j];
                o2[(long)a * D + b] = s2[i][j];
            }
        }
}

// out[i] = sum over the nc chunk partials of part[c * len + i], in chunk order.
__global__ void k_sum_chunks(const double *__restrict__ part, int nc, long len,
                             double *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int c = 0; c < nc; ++c) s += part[(long)c * len + i];
    out[i] = s;
}

}  // namespace

int dg_estep_fp64(dg_ctx *c, int isDA, double temp) {
    const double nbh = -c->beta / 2.0;
    const double it = isDA ? 1.0 / temp : 1.0;
    const int two = (isDA && it != 1.0) ? 1 : 0;
    const double logK = log((double)c->K);
    const int DJ = c->ldt / 16;
    if (DJ < 1 || DJ > 12) return dg_fail(c, DG_EINVAL, "FP64 path supports D <= 192");

    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[0], c->stream));
    const unsigned nt = (unsigned)((c->N + TP - 1) / TP);
    hipLaunchKernelGGL(k_e1_f64, dim3(nt), dim3(NT), 0, c->stream, c->T64, c->tt, c->N, c->Y64,
                       c->yy, c->K, c->ldt, nbh, it, two, isDA, logK, c->pm, c->pl, c->logg,
                       c->logh);
    DG_HIP(c, hipGetLastError());
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[1], c->stream));

    const long tiles = (c->N + TP - 1) / TP;
    const long chunk = ((tiles + c->NC - 1) / c->NC) * TP;
    dim3 grid((unsigned)((c->K + TK - 1) / TK), (unsigned)c->NC);
    switch (DJ) {
        case 1: launch_e2<1>(c, nbh, it, chunk, grid); break;
        case 2: launch_e2<2>(c, nbh, it, chunk, grid); break;
        case 3: launch_e2<3>(c, nbh, it, chunk, grid); break;
        case 4: launch_e2<4>(c, nbh, it, chunk, grid); break;
        case 5: launch_e2<5>(c, nbh, it, chunk, grid); break;
        case 6: launch_e2<6>(c, nbh, it, chunk, grid); break;
        case 7: launch_e2<7>(c, nbh, it, chunk, grid); break;
        case 8: launch_e2<8>(c, nbh, it, chunk, grid); break;
        case 9: launch_e2<9>(c, nbh, it, chunk, grid); break;
        case 10: launch_e2<10>(c, nbh, it, chunk, grid); break;
        case 11: launch_e2<11>(c, nbh, it, chunk, grid); break;
        default: launch_e2<12>(c, nbh, it, chunk, grid); break;
    }
    DG_HIP(c, hipGetLastError());
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[2], c->stream));

    const long KD = (long)c->K * c->D;
    const long nred = KD > c->K ? KD : c->K;
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((nred + 255) / 256)), dim3(256), 0,
                       c->stream, c->partRT, c->partG, c->NC, c->K, c->D, c->RT, c->G);
    DG_HIP(c, hipGetLastError());
    return dg_reduce_logs(c);
}

int dg_resp_dense_fp64(dg_ctx *c, double *dR, long ldR) {
    const double it = c->last_isDA ? 1.0 / c->last_temp : 1.0;
    dim3 grid((unsigned)((c->N + 255) / 256), (unsigned)c->K);
    hipLaunchKernelGGL(k_resp_dense, grid, dim3(256), 0, c->stream, c->T64, c->tt, c->N, c->Y64,
                       c->yy, c->K, c->ldt, -c->beta / 2.0, it, 1.0 / c->K, c->pm, c->pl, dR,
                       ldR);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_ctemp_moments_fp64(dg_ctx *c, int k0, int KB, int nc, double *dPart, double *dS) {
    const double it = c->last_isDA ? 1.0 / c->last_temp : 1.0;
    const int nt = (c->D + CM - 1) / CM;
    const long tiles = (c->N + CM - 1) / CM;
    const long chunk = ((tiles + nc - 1) / nc) * CM;
    dim3 grid((unsigned)(nt * (nt + 1) / 2), (unsigned)KB, (unsigned)nc);
    hipLaunchKernelGGL(k_ctemp_moments, grid, dim3(NT), 0, c->stream, c->T64, c->tt, c->N,
                       c->Y64, c->yy, k0, c->D, c->ldt, -c->beta / 2.0, it, 1.0 / c->K, c->pm,
                       c->pl, chunk, dPart);
    DG_HIP(c, hipGetLastError());
    if (nc > 1) {
        const long len = (long)KB * 2 * c->D * c->D;
        hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((len + 255) / 256)), dim3(256), 0,
                           c->stream, dPart, nc, len, dS);
        DG_HIP(c, hipGetLastError());
    }
    return DG_OK;
}

int dg_map_fp64(dg_ctx *c, const double *dX, int L, double *dM) {
    const double it = c->last_isDA ? 1.0 / c->last_temp : 1.0;
    hipLaunchKernelGGL(k_map_f64, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream,
                       c->T64, c->tt, c->N, c->Y64, c->yy, c->K, c->ldt, -c->beta / 2.0, it,
                       1.0 / c->K, c->pm, c->pl, dX, L, dM);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}
