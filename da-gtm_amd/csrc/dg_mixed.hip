// Mixed-precision E-step of DA-GTM on gfx950 (CDNA4 MFMA).
//
// Distances d = |y|^2 + |t|^2 - 2 y.t take their cross term from
// v_mfma_f32_32x32x16_f16 with FP32 accumulation.  Each centre coordinate is
// split into two FP16 terms y = (hi + lo) / S_y (22 significant bits) and
// the data either is exact in FP16 (binary fingerprints) or is split the same
// way; |y|^2 and |t|^2 are computed from the split values, so every distance
// is that of the represented points.  The exponent, the online log-sum-exp,
// the mask of exp underflow and every reduction run in FP32 (tile sums added
// with Kahan compensation); per-point log g / log h and everything after the
// E-step are FP64.
//
// Pass 1 (k_e1_mfma): a workgroup owns 128 points (32 per wave, their data
// operand held in registers) and streams all centres in 32-centre tiles
// through double-buffered LDS.  C[k][n] = Y T^T puts one point per lane
// column, so each lane reduces its 16 centre rows in registers and one
// cross-half shuffle finishes the tile.
//
// Pass 2 (k_e2_mfma): a workgroup owns 128 centres (32 per wave, centre
// operand in registers) and one chunk of points, streamed in 32-point blocks,
// two blocks in flight per wave.  C[n][k] = T Y^T leaves R in the accumulator
// layout that is exactly the B operand layout of the next product
// RT^T[d][k] += sum_n T^T[d][n] R[n][k], so R goes from the exponential
// straight into the second MFMA as an FP16 hi/lo pair without touching LDS.
// Each chunk writes FP32 partials that a deterministic FP64 pass sums.
#include <algorithm>
#include <type_traits>
#include <cmath>

#include "dg_internal.h"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int NTH = 256;          // threads per workgroup (4 waves)
constexpr int PT = 128;           // pass-1 points per workgroup
constexpr int CT = 128;           // pass-2 centres per workgroup
constexpr int KT = 32;            // pass-1 centres per tile
constexpr int BP = 32;            // pass-2 points per block
constexpr int PAD = 8;            // LDS row padding (halves)
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kZmin = -1075.0f;  // log2 of the FP64 exp underflow point
constexpr float kRLog2 = 14.0f;    // R is carried as R * 2^14 in FP16
constexpr float kMaskMargin = 64.0f;  // log2 distance beyond which masking is moot
constexpr double kLn2 = 0.6931471805599453;

__device__ __forceinline__ f16v mfma(h8 a, h8 b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }

__device__ __forceinline__ void kahan(float &s, float &c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---------------------------------------------------------------- pass 1
template <int KS, int TT>
__global__ __launch_bounds__(NTH) void k_e1_mfma(
    const _Float16 *__restrict__ T16, const float *__restrict__ tt32, long N, long Np,
    const _Float16 *__restrict__ Y16, const float *__restrict__ yy32, int K, int Kp,
    const double *__restrict__ mscal, float beta, float it, int two, int isDA, double logK,
    float *__restrict__ pB, float *__restrict__ pThr, float *__restrict__ blkThr,
    double *__restrict__ pm, double *__restrict__ logg, double *__restrict__ logh,
    double *__restrict__ partLog) {
    static_assert(PT == 4 * BP, "one pass-1 wave covers one pass-2 block");
    constexpr int ld = 16 * KS;
    constexpr int LR = ld + PAD;
    constexpr int ROW8 = ld / 8;                 // h8 chunks per row
    constexpr int CH = 2 * KT * ROW8;            // h8 chunks per tile (two terms)
    constexpr int PER = (CH + NTH - 1) / NTH;
    __shared__ __attribute__((aligned(16))) _Float16 Ys[2][2][KT][LR];
    __shared__ float aKs[2][KT];
    __shared__ double wsum[3][4];

    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5, j = l & 31;
    const long n = (long)blockIdx.x * PT + 32 * w + j;
    const float c1 = beta * kLog2e / (float)(mscal[0] * mscal[1]);
    const float cb = -0.5f * beta * kLog2e;
    // exponents are carried without this point's constant bn = cb |t_n|^2:
    // z' = z - bn; exp underflow (z < kZmin) is z' < thr
    const float bn = cb * tt32[n];
    const float thr = kZmin - bn;
    const float mmargin = kMaskMargin / it;

    h8 tb[TT][KS];
#pragma unroll
    for (int t = 0; t < TT; ++t)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            tb[t][ks] = *(const h8 *)(T16 + ((long)t * Np + n) * ld + 16 * ks + 8 * h);

    h8 pre[PER];
    float apre = 0.f;
    auto gload = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int c = tid + p * NTH;
            if (c < CH) {
                const int term = c / (KT * ROW8), rem = c % (KT * ROW8);
                const int row = rem / ROW8, col = rem % ROW8;
                pre[p] = *(const h8 *)(Y16 + ((long)term * Kp + k0 + row) * ld + col * 8);
            }
        }
        if (tid < KT) apre = (k0 + tid < K) ? cb * yy32[k0 + tid] : -INFINITY;
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int c = tid + p * NTH;
            if (c < CH) {
                const int term = c / (KT * ROW8), rem = c % (KT * ROW8);
                const int row = rem / ROW8, col = rem % ROW8;
                *(h8 *)&Ys[buf][term][row][col * 8] = pre[p];
            }
        }
        if (tid < KT) aKs[buf][tid] = apre;
    };

    float m = thr, S1 = 0.f, C1 = 0.f, S2 = 0.f, C2 = 0.f;
    const int ntiles = Kp / KT;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        if (t + 1 < ntiles) gload((t + 1) * KT);
        f16v acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const h8 a0 = *(const h8 *)&Ys[buf][0][j][16 * ks + 8 * h];
            const h8 a1 = *(const h8 *)&Ys[buf][1][j][16 * ks + 8 * h];
            acc = mfma(a0, tb[0][ks], acc);
            acc = mfma(a1, tb[0][ks], acc);
            if (TT == 2) acc = mfma(a0, tb[TT - 1][ks], acc);
        }
        float z[16];
        float zt = -INFINITY;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * h + (r & 3);
            z[r] = fmaf(c1, acc[r], aKs[buf][i]);
            zt = fmaxf(zt, z[r]);
        }
        zt = fmaxf(zt, __shfl_xor(zt, 32));
        // The underflow mask only matters when a masked element could weigh
        // more than 2^-64 of the running sum, i.e. when the tile maximum is
        // within kMaskMargin * temp of the threshold; otherwise skip it.
        if (!__all(zt - thr > mmargin)) {
            zt = -INFINITY;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                z[r] = (z[r] < thr) ? -INFINITY : z[r];
                zt = fmaxf(zt, z[r]);
            }
            zt = fmaxf(zt, __shfl_xor(zt, 32));
        }
        if (zt > m) {
            const float dm = m - zt;
            const float s1 = ex2(dm);
            S1 *= s1; C1 *= s1;
            if (two) {
                const float s2 = ex2(dm * it);
                S2 *= s2; C2 *= s2;
            }
            m = zt;
        }
        float t1 = 0.f, t2 = 0.f;
        const float nmit = -m * it;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            t1 += ex2(z[r] - m);
            if (two) t2 += ex2(fmaf(z[r], it, nmit));
        }
        kahan(S1, C1, t1);
        if (two) kahan(S2, C2, t2);
        if (t + 1 < ntiles) lstore(buf ^ 1);
        __syncthreads();
    }
    // lanes j and j+32 hold the same point (different centre rows) and share m
    const float S1o = __shfl_xor(S1, 32), C1o = __shfl_xor(C1, 32);
    const float S2o = __shfl_xor(S2, 32), C2o = __shfl_xor(C2, 32);
    double lg = 0.0, lh = 0.0, zc = 0.0;      // this lane's share of the workgroup sums
    // Padding points and zero columns get pB = -inf: every pass-2 exponent of
    // theirs is -inf, so R = 0 there without any masking (a fix-up adds 1/K
    // for zero columns afterwards).
    float pbv = -INFINITY, ptv = -INFINITY;
    if (h == 0) {
        const double s1 = ((double)S1 - C1) + ((double)S1o - C1o);
        const double s2 = two ? ((double)S2 - C2) + ((double)S2o - C2o) : s1;
        if (n >= N) {
        } else if (!(s1 > 0.0)) {          // zero column (mpicomm.c:382-390)
            lg = isDA ? -INFINITY : logK;
            lh = logK;
            zc = 1.0;
            pm[n] = -INFINITY;
            logg[n] = lg;
            logh[n] = lh;
        } else {
            const double L1 = log2(s1), L2 = log2(s2);
            const double ma = (double)m + (double)bn;        // the point's max exponent
            lg = kLn2 * (ma + L1);
            lh = kLn2 * (ma * it + L2);
            pm[n] = ma;
            logg[n] = lg;
            logh[n] = lh;
            pbv = (float)(-(double)it * m - L2 + kRLog2);
            ptv = (float)((double)it * ((double)thr - m) - L2 + kRLog2);
        }
        pB[n] = pbv;
        pThr[n] = ptv;
    }
    // This wave's 32 points are one pass-2 block: its largest mask threshold
    // tells pass 2 whether the block needs the underflow mask at all.
    float bt = ptv;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bt = fmaxf(bt, __shfl_xor(bt, o));
    if (l == 0) blkThr[(long)blockIdx.x * (PT / BP) + w] = bt;
    // workgroup partials of sum log g, sum log h and the zero-column count
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        lg += __shfl_xor(lg, o);
        lh += __shfl_xor(lh, o);
        zc += __shfl_xor(zc, o);
    }
    if (l == 0) { wsum[0][w] = lg; wsum[1][w] = lh; wsum[2][w] = zc; }
    __syncthreads();
    if (tid < 3)
        partLog[(long)blockIdx.x * 3 + tid] =
            (wsum[tid][0] + wsum[tid][1]) + (wsum[tid][2] + wsum[tid][3]);
}

// scal[0..2] = sums over the pass-1 workgroup partials (fixed tree order).
__global__ __launch_bounds__(1024) void k_sum_partlog(const double *__restrict__ part, long nb,
                                                      double *__restrict__ scal) {
    __shared__ double s[3][1024];
    double a = 0.0, b = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < nb; i += 1024) {
        a += part[3 * i]; b += part[3 * i + 1]; c += part[3 * i + 2];
    }
    s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = c;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
            s[2][threadIdx.x] += s[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[kScalSumLogG] = s[0][0];
        scal[kScalSumLogH] = s[1][0];
        scal[kScalNZero] = s[2][0];
    }
}

// ---------------------------------------------------------------- pass 2
// A workgroup owns 128 centres (32 per wave, both FP16 terms of Y held in
// registers as the B operand of the cross term) and one chunk of points,
// streamed in 32-point blocks through a three-slot LDS ring (the data rows
// for the cross term, the transposed rows for R T, pB and the mask
// thresholds).  Two blocks are in flight: while the cross-term MFMAs of block
// b+1 run, the wave turns block b's exponents into R (v_exp_f32, FP16 hi/lo
// pairs); while block b's R T MFMAs run, it forms block b+1's exponents.  Only
// VALU results (the exponents) cross the loop's back edge; every read of an
// MFMA result sits in the basic block of the MFMA that produced it.
//
// T16t stores the points of every 16-point group in the order
// 0-3, 8-11, 4-7, 12-15, so that one 16-byte LDS read gives a lane the eight
// points its R operand holds (the accumulator layout of the cross term).
// With GM, T16t row D is 1 for real points, and the R T product yields G in
// that row.  The 1-D grid is mapped so that the workgroups the hardware places
// on one XCD (linear id mod 8) share a point chunk and walk all its centre
// groups, keeping that chunk's data in the XCD's L2.
template <int KS, int TT, bool GM>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_e2_mfma(
    const _Float16 *__restrict__ T16, const _Float16 *__restrict__ T16t, long Np,
    const _Float16 *__restrict__ Y16, const float *__restrict__ yy32, int K, int Kp, int D,
    const double *__restrict__ mscal, float beta, float it,
    const float *__restrict__ pB, const float *__restrict__ pThr,
    const float *__restrict__ blkThr, long chunk, int KG,
    float *__restrict__ partRT, float *__restrict__ partG) {
    constexpr int ld = 16 * KS;
    constexpr int LR = ld + PAD;
    constexpr int DB = (ld + 31) / 32;
    constexpr int DR = 32 * DB;
    constexpr int TR = BP + PAD;
    constexpr int ROW8 = ld / 8;
    constexpr int CHA = TT * BP * ROW8;
    constexpr int CHB = TT * DR * (BP / 8);
    constexpr int PA = (CHA + NTH - 1) / NTH;
    constexpr int PB_ = (CHB + NTH - 1) / NTH;
    constexpr int NS = 3;                          // LDS ring slots
    __shared__ __attribute__((aligned(16))) _Float16 Ts[NS][TT][BP][LR];
    __shared__ __attribute__((aligned(16))) _Float16 Tts[NS][TT][DR][TR];
    __shared__ __attribute__((aligned(16))) float pBs[NS][BP];
    __shared__ __attribute__((aligned(16))) float pTs[NS][BP];
    const int L = blockIdx.x;
    const int xs = L & 7, sl = L >> 3;
    const int kg = sl % KG;
    const long cidx = (long)(sl / KG) * 8 + xs;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5, j = l & 31;
    const int k = kg * CT + 32 * w + j;
    const float c1 = beta * kLog2e / (float)(mscal[0] * mscal[1]);
    const float cb = -0.5f * beta * kLog2e;
    const float itc1 = it * c1;
    const float aKi = (k < K) ? it * (cb * yy32[k]) : -INFINITY;
    h8 yb[2][KS];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            yb[t][ks] = *(const h8 *)(Y16 + ((long)t * Kp + k) * ld + 16 * ks + 8 * h);
    f16v acc[DB];
#pragma unroll
    for (int b = 0; b < DB; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
    float G = 0.f, Gc = 0.f;                       // VALU row sums when !GM
    const long nb = cidx * chunk;
    const long ne = (nb + chunk < Np) ? nb + chunk : Np;
    const int nblocks = (ne > nb) ? (int)((ne - nb) / BP) : 0;

    h8 preA[PA], preB[PB_];
    float pp = 0.f;
    // global -> registers for block q (clamped into the chunk: a repeated
    // block is loaded but never used)
    auto gload = [&](int q) {
        const long n0 = nb + (long)(q < nblocks ? q : nblocks - 1) * BP;
#pragma unroll
        for (int p = 0; p < PA; ++p) {
            const int c = tid + p * NTH;
            if (c < CHA) {
                const int term = c / (BP * ROW8), rem = c % (BP * ROW8);
                const int row = rem / ROW8, col = rem % ROW8;
                preA[p] = *(const h8 *)(T16 + ((long)term * Np + n0 + row) * ld + col * 8);
            }
        }
#pragma unroll
        for (int p = 0; p < PB_; ++p) {
            const int c = tid + p * NTH;
            if (c < CHB) {
                const int term = c / (DR * (BP / 8)), rem = c % (DR * (BP / 8));
                const int row = rem / (BP / 8), col = rem % (BP / 8);
                preB[p] = *(const h8 *)(T16t + ((long)term * DR + row) * Np + n0 + col * 8);
            }
        }
        if (tid < BP) pp = pB[n0 + tid];
        else if (tid < 2 * BP) pp = pThr[n0 + tid - BP];
    };
    auto lstore = [&](int s) {
#pragma unroll
        for (int p = 0; p < PA; ++p) {
            const int c = tid + p * NTH;
            if (c < CHA) {
                const int term = c / (BP * ROW8), rem = c % (BP * ROW8);
                const int row = rem / ROW8, col = rem % ROW8;
                *(h8 *)&Ts[s][term][row][col * 8] = preA[p];
            }
        }
#pragma unroll
        for (int p = 0; p < PB_; ++p) {
            const int c = tid + p * NTH;
            if (c < CHB) {
                const int term = c / (DR * (BP / 8)), rem = c % (DR * (BP / 8));
                const int row = rem / (BP / 8), col = rem % (BP / 8);
                *(h8 *)&Tts[s][term][row][col * 8] = preB[p];
            }
        }
        if (tid < BP) pBs[s][tid] = pp;
        else if (tid < 2 * BP) pTs[s][tid - BP] = pp;
    };
    // cross term of the block in slot s -> exponents (base 2, R * 2^14 scale)
    auto cross_exp = [&](int s, float (&e)[16]) {
        f16v c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.f;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const h8 a0 = *(const h8 *)&Ts[s][0][j][16 * ks + 8 * h];
            c = mfma(a0, yb[0][ks], c);
            c = mfma(a0, yb[1][ks], c);
            if (TT == 2) {
                const h8 a1 = *(const h8 *)&Ts[s][TT - 1][j][16 * ks + 8 * h];
                c = mfma(a1, yb[0][ks], c);
            }
        }
#pragma unroll
        for (int m4 = 0; m4 < 4; ++m4) {
            const f4v pb = *(const f4v *)&pBs[s][8 * m4 + 4 * h];
#pragma unroll
            for (int q = 0; q < 4; ++q) e[4 * m4 + q] = fmaf(itc1, c[4 * m4 + q], aKi) + pb[q];
        }
    };
    // exp-before-pow underflow: elements below their point's threshold are 0
    auto mask = [&](int s, float (&e)[16]) {
#pragma unroll
        for (int m4 = 0; m4 < 4; ++m4) {
            const f4v pt = *(const f4v *)&pTs[s][8 * m4 + 4 * h];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                e[4 * m4 + q] = (e[4 * m4 + q] < pt[q]) ? -INFINITY : e[4 * m4 + q];
        }
    };
    // R = 2^e as FP16 hi/lo pairs in the B-operand order of the R T product
    auto resp = [&](const float (&e)[16], h8 (&rhi)[2], h8 (&rlo)[2]) {
        float gs = 0.f;
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const float R0 = ex2(e[r]), R1 = ex2(e[r + 1]);
            if (!GM) gs += R0 + R1;
            const h2 hp = {(_Float16)R0, (_Float16)R1};
            const h2 lp = {(_Float16)(R0 - (float)hp[0]), (_Float16)(R1 - (float)hp[1])};
            rhi[r >> 3][r & 7] = hp[0]; rhi[r >> 3][(r & 7) + 1] = hp[1];
            rlo[r >> 3][r & 7] = lp[0]; rlo[r >> 3][(r & 7) + 1] = lp[1];
        }
        if (!GM) kahan(G, Gc, gs);
    };
    auto rt = [&](int s, const h8 (&rhi)[2], const h8 (&rlo)[2]) {
#pragma unroll
        for (int db = 0; db < DB; ++db) {
#pragma unroll
            for (int sh = 0; sh < 2; ++sh) {
                const h8 a0 = *(const h8 *)&Tts[s][0][32 * db + j][16 * sh + 8 * h];
                acc[db] = mfma(a0, rhi[sh], acc[db]);
                acc[db] = mfma(a0, rlo[sh], acc[db]);
                if (TT == 2) {
                    const h8 a1 = *(const h8 *)&Tts[s][TT - 1][32 * db + j][16 * sh + 8 * h];
                    acc[db] = mfma(a1, rhi[sh], acc[db]);
                }
            }
        }
    };
    const long blk0 = nb / BP;                     // index of the chunk's first block

    if (nblocks > 0) {
        gload(0); lstore(0);
        gload(1); lstore(1);
        gload(2);
        __syncthreads();
        float e[16];
        cross_exp(0, e);
        if (blkThr[blk0] > -kMaskMargin) mask(0, e);
        for (int b = 0; b + 1 < nblocks; ++b) {
            const int s0 = b % NS, s1 = (b + 1) % NS, s2 = (b + 2) % NS;
            h8 rhi[2], rlo[2];
            resp(e, rhi, rlo);                     // block b: R from its exponents
            float en[16];
            cross_exp(s1, en);                     // block b+1: cross term
            rt(s0, rhi, rlo);                      // block b: R T
            lstore(s2);                            // block b+2 into the free slot
            gload(b + 3);
            if (blkThr[blk0 + b + 1] > -kMaskMargin) mask(s1, en);
#pragma unroll
            for (int r = 0; r < 16; ++r) e[r] = en[r];
            __syncthreads();
        }
        h8 rhi[2], rlo[2];
        resp(e, rhi, rlo);
        rt((nblocks - 1) % NS, rhi, rlo);
    }
    // the accumulators are read after the loop's exit branch: keep the
    // scheduler from moving those reads up and give the MFMA pipe time to drain
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
    const float inv = (float)(1.0 / (16384.0 * mscal[1]));
    const float ginv = 1.0f / 16384.0f;
    float gval = 0.f;
    if (k < K) {
#pragma unroll
        for (int db = 0; db < DB; ++db)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int d = 32 * db + 8 * (r >> 2) + 4 * h + (r & 3);
                if (d < D) partRT[(cidx * D + d) * K + k] = acc[db][r] * inv;
                if (GM && d == D) gval = acc[db][r];
            }
    }
    if (GM) {
        // row D lives in one half of the wave: take it from whichever holds it
        const float go = __shfl_xor(gval, 32);
        const bool mine = (((D & 31) >> 2) & 1) == h;
        if (h == 0 && k < K) partG[cidx * K + k] = (mine ? gval : go) * ginv;
    } else {
        const float Go = __shfl_xor(G, 32), Gco = __shfl_xor(Gc, 32);
        if (h == 0 && k < K) partG[cidx * K + k] = (((G - Gc) + (Go - Gco))) * ginv;
    }
}

// ---------------------------------------------------------------- helpers
__global__ void k_reduce32(const float *__restrict__ partRT, const float *__restrict__ partG,
                           int NC, int K, int D, double *__restrict__ RT,
                           double *__restrict__ G) {
    const long KD = (long)K * D;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;   // i = d * K + k
    if (i < KD) {
        const long d = i / K, k = i % K;
        double s = 0.0;
        for (int c = 0; c < NC; ++c) s += (double)partRT[(long)c * KD + i];
        RT[k * D + d] = s;
    }
    if (i < K) {
        double s = 0.0;
        for (int c = 0; c < NC; ++c) s += (double)partG[(long)c * K + i];
        G[i] = s;
    }
}

// Zero-column fall-back R = 1/K for the points pass 2 skipped: RT[k] +=
// (1/K) sum_{n in Z} t_n, G[k] += |Z| / K.  Exits at once when |Z| = 0.
__global__ __launch_bounds__(256) void k_zero_fix(const double *__restrict__ scal,
                                                  const double *__restrict__ T64, int ldt,
                                                  const double *__restrict__ pm, long N, int K,
                                                  int D, double *__restrict__ RT,
                                                  double *__restrict__ G) {
    const double nz = scal[kScalNZero];
    if (nz == 0.0) return;
    __shared__ double red[256];
    const int d = blockIdx.x;
    double s = 0.0;
    for (long n = threadIdx.x; n < N; n += 256)
        if (pm[n] == -INFINITY) s += T64[n * ldt + d];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double add = red[0] / K;
    for (int k = threadIdx.x; k < K; k += 256) {
        RT[(long)k * D + d] += add;
        if (d == 0) G[k] += nz / K;
    }
}

// Workgroup partials of max |x| and of the count of values not exact in
// FP16 over rows x D (stride ld): part[2 b], part[2 b + 1].
constexpr int kAmaxWG = 256;
__global__ __launch_bounds__(256) void k_absmax_part(const double *__restrict__ X, long rows,
                                                     int ld, int D, double *__restrict__ part) {
    __shared__ double smax[256], scnt[256];
    double mx = 0.0, cnt = 0.0;
    const long total = rows * D;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const double v = X[(e / D) * ld + (e % D)];
        mx = fmax(mx, fabs(v));
        const _Float16 hv = (_Float16)(float)v;
        cnt += ((double)(float)hv == v) ? 0.0 : 1.0;
    }
    smax[threadIdx.x] = mx; scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            scnt[threadIdx.x] += scnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = smax[0]; part[2 * blockIdx.x + 1] = scnt[0]; }
}

__global__ __launch_bounds__(256) void k_absmax_final(const double *__restrict__ part, int n,
                                                      double *__restrict__ out) {
    __shared__ double smax[256], scnt[256];
    double mx = 0.0, cnt = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) { mx = fmax(mx, part[2 * i]); cnt += part[2 * i + 1]; }
    smax[threadIdx.x] = mx; scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            scnt[threadIdx.x] += scnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = smax[0]; out[1] = scnt[0]; }
}

// Power-of-two scale putting max |x| at or below 2^14.
__device__ __forceinline__ double pow2_scale(double amax) {
    if (!(amax > 0.0)) return 1.0;
    const int e = ilogb(amax);                  // 2^e <= amax < 2^(e+1)
    return ldexp(1.0, 13 - e);
}

// Split T (N x D, stride ldt) into FP16 terms: T16[t][n][d], d < ld16, with
// rows n >= N and columns d >= D zero; tt32[n] = |t'|^2 of the split value.
__global__ void k_make_t16(const double *__restrict__ T64, int ldt, long N, long Np, int D,
                           int ld16, int TT, double scale, _Float16 *__restrict__ T16,
                           float *__restrict__ tt32) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Np) return;
    double ss = 0.0;
    for (int d = 0; d < ld16; ++d) {
        const double v = (n < N && d < D) ? T64[n * ldt + d] * scale : 0.0;
        const _Float16 hi = (_Float16)(float)v;
        const _Float16 lo = (TT == 2) ? (_Float16)(float)(v - (double)(float)hi) : (_Float16)0.0f;
        T16[n * ld16 + d] = hi;
        if (TT == 2) T16[(Np + n) * ld16 + d] = lo;
        const double rep = ((double)(float)hi + (double)(float)lo) / scale;
        ss = fma(rep, rep, ss);
    }
    tt32[n] = (float)ss;
}

// T16t[t][d][pos(n)] = T16[t][n][d] for d < DR (rows >= ld16 zero), with the
// points of each 16-point group stored in the order 0-3, 8-11, 4-7, 12-15
// (pos swaps the two middle quads) so that a lane's eight R-operand points
// are one 16-byte run.  With ones_row >= 0, row ones_row of term 0 is 1 for
// real points (n < N) and 0 for padding.  64 x 64 tiles.
__global__ void k_transpose16(const _Float16 *__restrict__ T16, long N, long Np, int ld16, int DR,
                              int ones_row, _Float16 *__restrict__ T16t) {
    __shared__ _Float16 tile[64][65];
    const long n0 = (long)blockIdx.x * 64;
    const int d0 = blockIdx.y * 64;
    const int t = blockIdx.z;
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        const int r = e / 64, c = e % 64;           // r: point, c: dim
        const long n = n0 + r;
        const int d = d0 + c;
        _Float16 v = (n < Np && d < ld16) ? T16[((long)t * Np + n) * ld16 + d] : (_Float16)0.0f;
        if (d == ones_row && t == 0) v = (n < N) ? (_Float16)1.0f : (_Float16)0.0f;
        tile[r][c] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        const int r = e / 64, c = e % 64;           // r: dim, c: stored position
        const int q = (c >> 2) & 3;
        const int src = (c & ~15) | ((((q & 1) << 1) | (q >> 1)) << 2) | (c & 3);
        const int d = d0 + r;
        const long n = n0 + c;
        if (d < DR && n < Np) T16t[((long)t * DR + d) * Np + n] = tile[src][r];
    }
}

// Split Y (K x D, stride ldt, FP64) into two FP16 terms with a power-of-two
// scale chosen from scal_in[0] = max |Y|: Y16[t][k][d], rows k >= K zero;
// yy32[k] = |y'|^2 of the split value; mscal[0] = the scale.
__global__ void k_make_y16(const double *__restrict__ Y64, int ldt, int K, int Kp, int D, int ld16,
                           const double *__restrict__ amax, _Float16 *__restrict__ Y16,
                           float *__restrict__ yy32, double *__restrict__ mscal) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const double scale = pow2_scale(amax[0]);
    if (k == 0) mscal[0] = scale;
    if (k >= Kp) return;
    double ss = 0.0;
    for (int d = 0; d < ld16; ++d) {
        const double v = (k < K && d < D) ? Y64[(long)k * ldt + d] * scale : 0.0;
        const _Float16 hi = (_Float16)(float)v;
        const _Float16 lo = (_Float16)(float)(v - (double)(float)hi);
        Y16[(long)k * ld16 + d] = hi;
        Y16[((long)Kp + k) * ld16 + d] = lo;
        const double rep = ((double)(float)hi + (double)(float)lo) / scale;
        ss = fma(rep, rep, ss);
    }
    yy32[k] = (float)ss;
}

// Supported numbers of 16-wide d steps (D is padded up to the next one).
constexpr int kKS[] = {1, 2, 3, 4, 6, 8, 11, 12};

int ks_for(int D) {
    const int need = (D + 15) / 16;
    for (int ks : kKS)
        if (ks >= need) return ks;
    return -1;
}

struct MixedBufs {
    _Float16 *T16t = nullptr;
    _Float16 *Y16 = nullptr;
    double *mscal = nullptr;
    double *partLog = nullptr;    // 3 per pass-1 workgroup
    float *blkThr = nullptr;      // per 32-point block: largest mask threshold
    double *amaxPart = nullptr;   // 2 x kAmaxWG
    long Np = 0;
    int Kp = 0, KS = 0, TT = 1, NC = 1;
};

MixedBufs *mb(dg_ctx *c) { return reinterpret_cast<MixedBufs *>(c->mix); }

template <int KS, int TT>
void launch_mixed(dg_ctx *c, MixedBufs *b, float beta, float it, int two, int isDA,
                  double logK) {
    hipLaunchKernelGGL((k_e1_mfma<KS, TT>), dim3((unsigned)(b->Np / PT)), dim3(NTH), 0, c->stream,
                       (const _Float16 *)c->T16, c->tt32, c->N, b->Np, b->Y16, c->yy32, c->K,
                       b->Kp, b->mscal, beta, it, two, isDA, logK, c->pm32, c->pl32, b->blkThr,
                       c->pm, c->logg, c->logh, b->partLog);
    if (c->timing) (void)hipEventRecord(c->ev[1], c->stream);
    const long blocks = b->Np / BP;
    const long chunk = ((blocks + b->NC - 1) / b->NC) * BP;
    const int KG = b->Kp / CT;
    // G comes out of the R T product when the padded d range has a spare row
    constexpr int DR = 32 * ((16 * KS + 31) / 32);
    if (DR > c->D)
        hipLaunchKernelGGL((k_e2_mfma<KS, TT, true>), dim3((unsigned)(KG * b->NC)), dim3(NTH), 0,
                           c->stream, (const _Float16 *)c->T16, b->T16t, b->Np, b->Y16, c->yy32,
                           c->K, b->Kp, c->D, b->mscal, beta, it, c->pm32, c->pl32, b->blkThr,
                           chunk, KG, c->partRT32, c->partG32);
    else
        hipLaunchKernelGGL((k_e2_mfma<KS, TT, false>), dim3((unsigned)(KG * b->NC)), dim3(NTH), 0,
                           c->stream, (const _Float16 *)c->T16, b->T16t, b->Np, b->Y16, c->yy32,
                           c->K, b->Kp, c->D, b->mscal, beta, it, c->pm32, c->pl32, b->blkThr,
                           chunk, KG, c->partRT32, c->partG32);
}

template <int TT>
int dispatch_ks(dg_ctx *c, MixedBufs *b, float beta, float it, int two, int isDA, double logK) {
    switch (b->KS) {
        case 1: launch_mixed<1, TT>(c, b, beta, it, two, isDA, logK); break;
        case 2: launch_mixed<2, TT>(c, b, beta, it, two, isDA, logK); break;
        case 3: launch_mixed<3, TT>(c, b, beta, it, two, isDA, logK); break;
        case 4: launch_mixed<4, TT>(c, b, beta, it, two, isDA, logK); break;
        case 6: launch_mixed<6, TT>(c, b, beta, it, two, isDA, logK); break;
        case 8: launch_mixed<8, TT>(c, b, beta, it, two, isDA, logK); break;
        case 11:
        case 12:
            // the split data's pass-2 LDS ring (three slots, two terms) exceeds 160 KB above D = 128
            if constexpr (TT == 1) {
                if (b->KS == 11) launch_mixed<11, 1>(c, b, beta, it, two, isDA, logK);
                else launch_mixed<12, 1>(c, b, beta, it, two, isDA, logK);
                break;
            } else {
                return dg_fail(c, DG_EINVAL,
                               "mixed path: data not exact in FP16 is limited to D <= 128");
            }
        default: return dg_fail(c, DG_EINVAL, "mixed path: unsupported D");
    }
    return DG_OK;
}

template <typename T>
int dalloc(dg_ctx *c, T **p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return dg_fail(c, DG_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return DG_OK;
}

}  // namespace

void dg_mixed_free(dg_ctx *c) {
    MixedBufs *b = mb(c);
    if (b) {
        if (b->T16t) (void)hipFree(b->T16t);
        if (b->Y16) (void)hipFree(b->Y16);
        if (b->mscal) (void)hipFree(b->mscal);
        if (b->partLog) (void)hipFree(b->partLog);
        if (b->blkThr) (void)hipFree(b->blkThr);
        if (b->amaxPart) (void)hipFree(b->amaxPart);
        delete b;
    }
    c->mix = nullptr;
    if (c->T16) (void)hipFree(c->T16);
    if (c->tt32) (void)hipFree(c->tt32);
    if (c->yy32) (void)hipFree(c->yy32);
    if (c->pm32) (void)hipFree(c->pm32);
    if (c->pl32) (void)hipFree(c->pl32);
    if (c->partRT32) (void)hipFree(c->partRT32);
    if (c->partG32) (void)hipFree(c->partG32);
    c->T16 = nullptr; c->tt32 = nullptr; c->yy32 = nullptr; c->pm32 = nullptr;
    c->pl32 = nullptr; c->partRT32 = nullptr; c->partG32 = nullptr;
    c->mixed_ready = 0;
}

int dg_mixed_prepare_t(dg_ctx *c) {
    dg_mixed_free(c);
    const int KS = ks_for(c->D);
    if (KS < 0) return dg_fail(c, DG_EINVAL, "mixed path supports D <= 192");
    MixedBufs *b = new (std::nothrow) MixedBufs();
    if (!b) return dg_fail(c, DG_ENOMEM, "host allocation");
    c->mix = b;
    b->KS = KS;
    const int ld16 = 16 * KS;
    const int DR = 32 * ((ld16 + 31) / 32);
    c->ldh = ld16;
    b->Np = ((c->N + PT - 1) / PT) * PT;
    b->Kp = ((c->K + CT - 1) / CT) * CT;
    // chunks of pass 2: ~8K points each, enough workgroups to fill the chip
    const long blocks = b->Np / BP;
    long nc = (c->N + 8191) / 8192;
    const long want = (2048 + (b->Kp / CT) - 1) / (b->Kp / CT);
    if (nc < want) nc = want;
    if (nc > blocks) nc = blocks;
    if (nc < 1) nc = 1;
    nc = ((nc + 7) / 8) * 8;          // the pass-2 grid maps 8 chunks per XCD round
    b->NC = (int)nc;

    int rc;
    if ((rc = dalloc(c, &b->amaxPart, (size_t)2 * kAmaxWG))) return rc;
    // FP16-exactness and range of T
    hipLaunchKernelGGL(k_absmax_part, dim3(kAmaxWG), dim3(256), 0, c->stream, c->T64, c->N, c->ldt,
                       c->D, b->amaxPart);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(256), 0, c->stream, b->amaxPart, kAmaxWG,
                       c->scal + 6);
    DG_HIP(c, hipGetLastError());
    DG_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, kScalCount * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    DG_HIP(c, hipStreamSynchronize(c->stream));
    const double amax = c->h_scal[6], inexact = c->h_scal[7];
    c->t_exact = (inexact == 0.0 && amax <= 2048.0) ? 1 : 0;
    b->TT = c->t_exact ? 1 : 2;
    if (b->TT == 2 && KS > 8)
        return dg_fail(c, DG_EINVAL,
                       "mixed path: data not exact in FP16 is limited to D <= 128 "
                       "(use DG_FP64, or FP16-exact data such as binary fingerprints)");
    double tscale = 1.0;
    if (!c->t_exact && amax > 0.0) tscale = ldexp(1.0, 13 - ilogb(amax));
    c->t_scale = (float)tscale;

    if ((rc = dalloc(c, (_Float16 **)&c->T16, (size_t)b->TT * b->Np * ld16)) ||
        (rc = dalloc(c, &b->T16t, (size_t)b->TT * DR * b->Np)) ||
        (rc = dalloc(c, &c->tt32, (size_t)b->Np)) ||
        (rc = dalloc(c, &b->Y16, (size_t)2 * b->Kp * ld16)) ||
        (rc = dalloc(c, &c->yy32, (size_t)b->Kp)) || (rc = dalloc(c, &b->mscal, 4)) ||
        (rc = dalloc(c, &b->partLog, (size_t)3 * (b->Np / PT))) ||
        (rc = dalloc(c, &b->blkThr, (size_t)(b->Np / BP))) ||
        (rc = dalloc(c, &c->pm32, (size_t)b->Np)) || (rc = dalloc(c, &c->pl32, (size_t)b->Np)) ||
        (rc = dalloc(c, &c->partRT32, (size_t)b->NC * c->K * c->D)) ||
        (rc = dalloc(c, &c->partG32, (size_t)b->NC * c->K)))
        return rc;
    double ms[4] = {1.0, tscale, 0.0, 0.0};
    DG_HIP(c, hipMemcpyAsync(b->mscal, ms, sizeof(ms), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_make_t16, dim3((unsigned)((b->Np + 255) / 256)), dim3(256), 0, c->stream,
                       c->T64, c->ldt, c->N, b->Np, c->D, ld16, b->TT, tscale,
                       (_Float16 *)c->T16, c->tt32);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_transpose16,
                       dim3((unsigned)((b->Np + 63) / 64), (unsigned)((DR + 63) / 64),
                            (unsigned)b->TT),
                       dim3(256), 0, c->stream, (const _Float16 *)c->T16, c->N, b->Np, ld16, DR,
                       DR > c->D ? c->D : -1, b->T16t);
    DG_HIP(c, hipGetLastError());
    DG_HIP(c, hipStreamSynchronize(c->stream));
    c->mixed_ready = 1;
    return DG_OK;
}

int dg_mixed_prepare_y(dg_ctx *c) {
    MixedBufs *b = mb(c);
    if (!b || !c->mixed_ready) return dg_fail(c, DG_ESTATE, "mixed buffers not prepared");
    const int nwg = (int)std::min<long>(kAmaxWG, ((long)c->K * c->D + 4095) / 4096);
    hipLaunchKernelGGL(k_absmax_part, dim3((unsigned)nwg), dim3(256), 0, c->stream, c->Y64,
                       (long)c->K, c->ldt, c->D, b->amaxPart);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(256), 0, c->stream, b->amaxPart, nwg,
                       c->scal + 6);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_make_y16, dim3((unsigned)((b->Kp + 255) / 256)), dim3(256), 0, c->stream,
                       c->Y64, c->ldt, c->K, b->Kp, c->D, c->ldh, c->scal + 6, b->Y16, c->yy32,
                       b->mscal);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

int dg_estep_mixed(dg_ctx *c, int isDA, double temp) {
    MixedBufs *b = mb(c);
    if (!b || !c->mixed_ready) return dg_fail(c, DG_ESTATE, "mixed buffers not prepared");
    const float it = isDA ? (float)(1.0 / temp) : 1.0f;
    const int two = (isDA && it != 1.0f) ? 1 : 0;
    const double logK = log((double)c->K);
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[0], c->stream));
    int rc = (b->TT == 1) ? dispatch_ks<1>(c, b, (float)c->beta, it, two, isDA, logK)
                          : dispatch_ks<2>(c, b, (float)c->beta, it, two, isDA, logK);
    if (rc) return rc;
    DG_HIP(c, hipGetLastError());
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[2], c->stream));
    const long KD = (long)c->K * c->D;
    const long nred = KD > c->K ? KD : c->K;
    hipLaunchKernelGGL(k_reduce32, dim3((unsigned)((nred + 255) / 256)), dim3(256), 0, c->stream,
                       c->partRT32, c->partG32, b->NC, c->K, c->D, c->RT, c->G);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_sum_partlog, dim3(1), dim3(1024), 0, c->stream, b->partLog, b->Np / PT,
                       c->scal);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_zero_fix, dim3((unsigned)c->D), dim3(256), 0, c->stream, c->scal, c->T64,
                       c->ldt, c->pm, c->N, c->K, c->D, c->RT, c->G);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}
