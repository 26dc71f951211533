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
// through an LDS ring, two tiles in flight per wave.  C[k][n] = Y T^T puts
// one point per lane column, so each lane reduces its 16 centre rows in
// registers and one cross-half exchange finishes the tile.
//
// Pass 2 (k_e2_mfma): a workgroup owns 128 centres (32 per wave, centre
// operand in registers) and one chunk of points, streamed in 32-point blocks,
// two blocks in flight per wave.  C[n][k] = T Y^T leaves R in the accumulator
// layout that is exactly the B operand layout of the next product
// RT^T[d][k] += sum_n T^T[d][n] R[n][k], so R goes from the exponential
// straight into the second MFMA as an FP16 hi/lo pair without touching LDS.
// Each chunk writes FP32 partials that a deterministic FP64 pass sums.
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <cmath>

#include "dg_internal.h"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int NTH = 256;          // threads per workgroup (4 waves)
constexpr int PT = 128;           // pass-1 points per workgroup
constexpr int CT = 128;           // pass-2 centres per workgroup
constexpr int KT = 32;            // pass-1 centres per tile
constexpr int BP = 32;            // pass-2 points per block
constexpr int AKS = KT + 4;       // pass-1 aK per tile: 32 values, the largest, 3 pads
constexpr int PAD = 8;            // LDS row padding (halves)
// Pass 2 visits, per (point chunk, centre group) pair, at most MAXR runs of
// adjacent blocks; dead gaps of at most RGAP blocks stay inside a run (none:
// a run's restart costs less than a dead block).  The pair's entry in the run
// table is RS int2: {runs, blocks}, then the runs {first block within the
// chunk, count}.
constexpr int MAXR = 31;
constexpr int RS = MAXR + 1;
constexpr int RGAP = 0;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kZmin = -1075.0f;  // log2 of the FP64 exp underflow point
constexpr float kRLog2 = 14.0f;    // R is carried as R * 2^14 in FP16
constexpr float kMaskMargin = 64.0f;  // log2 distance beyond which masking is moot
constexpr double kLn2 = 0.6931471805599453;

__device__ __forceinline__ f16v mfma(h8 a, h8 b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }

// max of x over lanes l and l ^ 32.  v_permlane32_swap exchanges the upper
// half of its first operand with the lower half of its second; with x in both,
// every lane ends up holding {x_l, x_(l^32)} in the two results.  A VALU
// instruction, unlike ds_bpermute, which goes through the LDS pipeline.
__device__ __forceinline__ float xhalf_max(float x) {
    const unsigned u = __builtin_bit_cast(unsigned, x);
    const auto p = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return fmaxf(__builtin_bit_cast(float, (unsigned)p[0]), __builtin_bit_cast(float, (unsigned)p[1]));
}

__device__ __forceinline__ void kahan(float &s, float &c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---------------------------------------------------------------- pass 1
// A workgroup owns 128 points (32 per wave; their FP16 data held in
// registers as the B operand) and streams all centres in 32-centre tiles
// through a two-slot LDS ring filled by LDS-DMA (both FP16 terms of the
// tile's centres and their exponent terms aK = -beta log2(e) |y|^2 / 2).
// C[k][n] = Y T^T puts one point per lane column, so each lane reduces its 16
// centre rows in registers and one cross-half exchange finishes the tile.
// Two tiles are in flight per wave, in a fixed issue order held by
// sched_barrier: each cross-term MFMA of tile t+1 is followed by the
// exponentials and sums of one centre row of tile t (or a DMA load of tile
// t+2); tile t+1's exponents, maximum and rescaling then follow in the basic
// block of its MFMAs.  Only VALU results (the exponents) cross the loop's back
// edge.  The register cap of two waves per SIMD (the kernel needs at most 146
// registers and runs three) makes the compiler keep the accumulator in
// ordinary VGPRs, where the exponent FMAs read it directly; in AGPRs every
// element would first need a v_accvgpr_read.
// The body of one workgroup (point tile bx); k_e1_mfma runs one per
// workgroup, kb_e1_mfma takes tiles of many instances from a work queue.
// GMX: this instantiation keeps the centre-group maxima (gmx) and the pass-2
// bounds (bnd); without it both pointers are ignored.
// SCR (data exact in FP16, GMX): a tile's cross term takes the hi term of Y
// first; when that product plus a bound on the lo term puts every element of
// the wave's tile more than `smargin` below its point's running maximum, the
// lo term and the tile's exponentials are left out (each such element weighs
// less than 2^-48 of its point's sums) and the bound enters the group maximum.
// t0: the tile the sweep starts from (a multiple of four, so that groups stay
// aligned); the sweep wraps around.
template <int KS, int TT, bool TWO, bool GMX = true, bool SCR = false>
__device__ __forceinline__ void e1_body(
    const long bx, const int tid, const _Float16 *__restrict__ T16, const float *__restrict__ tt32, long N,
    long Np, const _Float16 *__restrict__ Y16, const float *__restrict__ aK32, int Kp,
    const double *__restrict__ mscal, float beta, float it, int isDA, double logK,
    float *__restrict__ pP, float *__restrict__ blkThr, double *__restrict__ pm,
    double *__restrict__ logg, double *__restrict__ logh, double *__restrict__ partLog,
    const int *__restrict__ perm, float *__restrict__ gmx, float *__restrict__ bnd,
    const int t0 = 0, const float smargin = INFINITY) {
    static_assert(!SCR || (TT == 1 && GMX), "screening needs exact data and the group maxima");
    static_assert(PT == 4 * BP, "one pass-1 wave covers one pass-2 block");
    static_assert(CT == 4 * KT, "four pass-1 tiles make one pass-2 centre group");
    constexpr int ld = 16 * KS;
    constexpr int LR = ld + PAD;                   // centre row in LDS (halves), padding last
    constexpr int RC = LR / 8;                     // 16-byte chunks per row
    constexpr int NIA = (2 * KT * RC + 63) / 64;   // 1 KB LDS-DMA loads: both Y terms,
    constexpr int NI = NIA + 1;                    // + the tile's aK
    constexpr int NIW = (NI + 3) / 4;              // per wave
    constexpr int SLOT = 4 * NIW * 512;            // halves per ring slot
    constexpr int OFF_AK = NIA * 512;
    // Iteration t reads only the slot of tile t+1 (tile t's exponents are in
    // registers), so the DMA of tile t+2 goes to the slot tile t came from:
    // two slots suffice.
    constexpr int NS = 2;
    constexpr int NX = (TT + 1) * KS;              // cross-term MFMAs per tile
    static_assert(NS * SLOT * 2 <= 80 * 1024, "two workgroups per CU");
    // one object per ring slot: with the slot a constant in each unrolled
    // iteration, the compiler sees DMA targets and operand reads as distinct
    __shared__ __attribute__((aligned(1024))) _Float16 smem0[SLOT];
    __shared__ __attribute__((aligned(1024))) _Float16 smem1[SLOT];
    static_assert(NS == 2, "two slot objects");
    auto sp = [&](int s) -> _Float16 * { return s ? smem1 : smem0; };
    __shared__ double wsum[3][4];

    // (the & 3 bounds the wave index for the compiler when tid is opaque, as in
    // the persistent kernels: it then still proves each LDS-DMA target disjoint
    // from the LDS reads of the other ring slots and waits for no DMA there)
    const int w = (tid >> 6) & 3, l = tid & 63, h = l >> 5, j = l & 31;
    const int wu = __builtin_amdgcn_readfirstlane(w) & 3;
    const long n = bx * PT + 32 * w + j;
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
    // SCR: the lo term's share of an exponent is at most c1 * 8 * sum_d |t'_d|
    // (the split leaves |ylo'| <= 8 for values below 2^14), plus one unit of
    // slack for the FP32 rounding of the hi-term product
    float errz = 0.f;
    if (SCR) {
        float a = 0.f;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) a += fabsf((float)tb[0][ks][e]);
        a += __shfl_xor(a, 32);                    // the other half of d
        errz = c1 * 8.0f * a * 1.001f + 1.0f;
    }

    // DMA: load I of a tile fills chunks 64 I .. 64 I + 63 of the slot in lane
    // order ([Y hi rows | Y lo rows | aK]); wave w issues loads w, w + 4, ...
    // Padding chunks take a copy of the load's first source chunk.
    unsigned voff[NIW];
    const char *dbase[NIW];
    unsigned dstep[NIW];
#pragma unroll
    for (int i = 0; i < NIW; ++i) {
        const int I = wu + 4 * i;
        unsigned o = 0;
        if (I < NIA) {
            const int g = I * 64 + l;
            const int t = g / (KT * RC), rem = g % (KT * RC), row = rem / RC, col = rem % RC;
            if (t < 2 && col < ld / 8) o = (unsigned)((((long)t * Kp + row) * ld + col * 8) * 2);
        } else if (I == NIA) {
            if (l < AKS / 4) o = (unsigned)(l * 16);   // the tile's aK, its largest, pads
        }
        voff[i] = o;
        dbase[i] = (I < NIA) ? (const char *)Y16 : (const char *)aK32;
        dstep[i] = (I < NIA) ? (unsigned)(KT * ld * 2) : (unsigned)(AKS * 4);
    }
    const int ntiles = Kp / KT;
    // loop position q holds tile q + t0 (mod ntiles)
    auto tile_of = [&](int q) { return (q + t0 < ntiles) ? q + t0 : q + t0 - ntiles; };
    auto dma = [&](int q, int s, int i) {
        const unsigned qc = (unsigned)tile_of(q < ntiles ? q : ntiles - 1);
        const int I = wu + 4 * i;
        __builtin_amdgcn_global_load_lds(
            (const void *)(dbase[i] + qc * dstep[i] + voff[i]),
            (__attribute__((address_space(3))) void *)&sp(s)[I * 512], 16, 0, 0);
    };
    // Marks the lane offsets as redefined once per loop iteration (an empty
    // asm, no instruction).  Otherwise the compiler hoists base + offset into
    // a 64-bit VGPR pair per load and adds the block step with a VALU
    // v_lshl_add_u64 before every DMA; with the offsets opaque, the address
    // stays a scalar base (SGPR pair, SALU arithmetic) plus a 32-bit lane
    // offset, the saddr form of global_load_lds.
    auto dma_offsets_opaque = [&]() {
#pragma unroll
        for (int i = 0; i < NIW; ++i) asm volatile("" : "+v"(voff[i]));
    };
    auto ldY = [&](int s, int t, int ks) {         // A operand: centre row j, term t
        return *(const h8 *)&sp(s)[(t * KT + j) * LR + 16 * ks + 8 * h];
    };
    auto ldK = [&](int s, int m4) {                // aK of rows 8 m4 + 4 h .. + 3
        return *(const f4v *)&sp(s)[OFF_AK + 2 * (8 * m4 + 4 * h)];
    };
    auto ldKmax = [&](int s) {                     // the tile's largest aK
        return *(const float *)&sp(s)[OFF_AK + 2 * KT];
    };
    // cross-term MFMA number m (m = (TT + 1) ks + which) of the tile whose
    // Y rows are a0 (hi) / a1 (lo)
    // (with the screen, every hi-term MFMA comes first and the lo terms after
    // them: pass 2 accumulates its cross term in the same order, so both
    // passes form bit-identical exponents)
    auto xks = [&](int m) { return SCR ? (m < KS ? m : m - KS) : m / (TT + 1); };
    auto xmfma = [&](int m, const h8 &a0, const h8 &a1, f16v &c) {
        const int ks = xks(m), wq = SCR ? (m < KS ? 0 : 1) : m % (TT + 1);
        if (wq == 0) c = mfma(a0, tb[0][ks], c);
        else if (wq == 1) c = mfma(a1, tb[0][ks], c);
        else c = mfma(a0, tb[TT - 1][ks], c);
    };

    float m = thr, S1 = 0.f, C1 = 0.f, S2 = 0.f, C2 = 0.f;
    float t1 = 0.f, t2 = 0.f, nmit = 0.f;
    // the maximum over the current centre group (four tiles)
    float gz = -INFINITY;
    // exponents of a tile from its accumulator, the tile maximum, the
    // underflow mask when it can matter, and the online rescaling
    auto tile_exp = [&](const f4v (&akr)[4], const f16v &c, float (&z)[16]) {
        float zt = -INFINITY;
#pragma unroll
        for (int m4 = 0; m4 < 4; ++m4) {
            const f4v ak = akr[m4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                z[4 * m4 + q] = fmaf(c1, c[4 * m4 + q], ak[q]);
                zt = fmaxf(zt, z[4 * m4 + q]);
            }
        }
        zt = xhalf_max(zt);
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
            zt = xhalf_max(zt);
        }
        if (GMX) gz = fmaxf(gz, zt);
        if (zt > m) {
            const float dm = m - zt;
            const float s1 = ex2(dm);
            S1 *= s1; C1 *= s1;
            if (TWO) {
                const float s2 = ex2(dm * it);
                S2 *= s2; C2 *= s2;
            }
            m = zt;
        }
        nmit = -m * it;
    };
    // after the last tile of a centre group: its maximum per point (lanes of
    // both halves hold the same value; the first half stores it)
    // (a scalar base for this wave's 32 points plus the lane offset)
    auto group_end = [&](const int tix) {
        if (GMX && gmx && (tix & 3) == 3) {
            float *gp = gmx + (long)(tile_of(tix) >> 2) * Np + bx * PT + 32 * wu;
            if (h == 0) gp[j] = gz;
            gz = -INFINITY;
        }
    };
    // exponentials and sums of a tile, in steps: step 2r adds centre row r to
    // the first sum, step 2r+1 to the second (nothing when !TWO); row 0
    // starts the sums (an add to 0.0f is not folded away: -0 + 0 = +0)
    constexpr int NRW = 32;
    auto row_sum = [&](const float (&z)[16], int st) {
        const int r = st >> 1;
        if ((st & 1) == 0) {
            const float v = ex2(z[r] - m);
            t1 = (r == 0) ? v : t1 + v;
        } else if (TWO) {
            const float v = ex2(fmaf(z[r], it, nmit));
            t2 = (r == 0) ? v : t2 + v;
        }
    };
    // The empty asm pins the sums at this point: without it the compiler
    // sinks the whole chain of exponentials below the next branch, out of the
    // MFMA block they are meant to overlap.
    auto tile_close = [&]() {
        kahan(S1, C1, t1);
        if (TWO) kahan(S2, C2, t2);
        asm volatile("" : "+v"(S1), "+v"(C1), "+v"(S2), "+v"(C2));
    };

#pragma unroll
    for (int i = 0; i < NIW; ++i) dma(0, 0, i);
#pragma unroll
    for (int i = 0; i < NIW; ++i) dma(1, 1, i);
    __syncthreads();
    float z[16];
    {
        f16v c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.f;
#pragma unroll
        for (int mm = 0; mm < NX; ++mm) {
            const int ks = xks(mm);
            xmfma(mm, ldY(0, 0, ks), ldY(0, 1, ks), c);
        }
        f4v ak0[4];
#pragma unroll
        for (int m4 = 0; m4 < 4; ++m4) ak0[m4] = ldK(0, m4);
        tile_exp(ak0, c, z);
    }
    // iteration 0 refills tile 0's slot: every wave must be done reading it
    __syncthreads();
    // work steps spread evenly over the MFMA slots: the row sums of tile t,
    // then the DMA loads (slot m runs steps m W1 / NX .. (m + 1) W1 / NX - 1)
    constexpr int W1 = NRW + NIW;
    // iteration t; the slots are template constants, the loop below runs two
    // iterations per trip, one per slot rotation
    auto iter1 = [&](const int t, auto S1c) {
        constexpr int s1 = decltype(S1c)::value, s2 = (s1 + 1) % NS;
        dma_offsets_opaque();
        f16v c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.f;
        f4v akr[4];                                // tile t+1's aK
        h8 ya0[KS], ya1[KS];
#pragma unroll
        for (int ks = 0; ks < 2 && ks < KS; ++ks) { ya0[ks] = ldY(s1, 0, ks); ya1[ks] = ldY(s1, 1, ks); }
#pragma unroll
        for (int mm = 0; mm < NX; ++mm) {
            const int ks = mm / (TT + 1);
            if (mm % (TT + 1) == 0 && ks + 2 < KS) {
                ya0[ks + 2] = ldY(s1, 0, ks + 2);
                ya1[ks + 2] = ldY(s1, 1, ks + 2);
            }
            xmfma(mm, ya0[ks], ya1[ks], c);
            // tile t+1's aK goes to registers in the slot before the first DMA:
            // after an LDS-DMA the compiler puts vmcnt(0) (the DMA's completion)
            // before any LDS read it cannot prove disjoint from the DMA target
            if (mm == NRW * NX / W1 - 1) {
#pragma unroll
                for (int m4 = 0; m4 < 4; ++m4) akr[m4] = ldK(s1, m4);
            }
#pragma unroll
            for (int st = mm * W1 / NX; st < (mm + 1) * W1 / NX; ++st) {
                if (st < NRW) row_sum(z, st);
                else dma(t + 2, s2, st - NRW);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        tile_close();
        tile_exp(akr, c, z);
        group_end(t + 1);
        __syncthreads();                           // also waits for this iteration's DMA
    };
    // SCR iteration t: phase A, tile t+1's hi-term MFMAs with the DMA loads
    // of tile t+2; the screen; then, unless tile t+1 is screened out, phase B
    // (its lo-term MFMAs), its exponents and its row sums, all in this
    // iteration: a screened tile costs its hi-term MFMAs and the screen only.
    // (The row sums no longer overlap the next tile's MFMAs within a wave;
    // the other waves of the SIMD fill that, and the per-tile branch around
    // them costs less than the interleave did: 3.81 against 4.06 ms.)
    // Returns whether tile t+1 is screened out.
    auto iter1s = [&](const int t, auto S1c) -> bool {
        constexpr int s1 = decltype(S1c)::value, s2 = (s1 + 1) % NS;
        f16v c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.f;
        const float akmax = ldKmax(s1);
        h8 ya0[KS], ya1[KS];
#pragma unroll
        for (int ks = 0; ks < 2 && ks < KS; ++ks) ya0[ks] = ldY(s1, 0, ks);
        dma_offsets_opaque();
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (ks + 2 < KS) ya0[ks + 2] = ldY(s1, 0, ks + 2);
            if (ks >= KS - 2) ya1[ks - (KS - 2)] = ldY(s1, 1, ks - (KS - 2));
            c = mfma(ya0[ks], tb[0][ks], c);
#pragma unroll
            for (int st = ks * NIW / KS; st < (ks + 1) * NIW / KS; ++st) dma(t + 2, s2, st);
            __builtin_amdgcn_sched_barrier(0);
        }
        // the screen: an upper bound of every exponent of tile t+1 per point
        // (c1 > 0: the largest product, the tile's largest aK, the lo-term bound)
        float cm = c[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) cm = fmaxf(cm, c[r]);
        const float ub = fmaf(c1, xhalf_max(cm), akmax) + errz;
        const bool out = __all(ub < m - smargin);
        if (!out) {                                // phase B, the exponents and the sums
            f4v akr[4];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                if (ks + 2 < KS) ya1[ks + 2] = ldY(s1, 1, ks + 2);
                if (ks == KS - 1) {                // tile t+1's aK, for its exponents
#pragma unroll
                    for (int m4 = 0; m4 < 4; ++m4) akr[m4] = ldK(s1, m4);
                }
                c = mfma(ya1[ks], tb[0][ks], c);
                __builtin_amdgcn_sched_barrier(0);
            }
            tile_exp(akr, c, z);
#pragma unroll
            for (int st = 0; st < NRW; ++st) row_sum(z, st);
            tile_close();
        } else if (GMX) {
            gz = fmaxf(gz, ub);
        }
        group_end(t + 1);
        __syncthreads();                           // also waits for this iteration's DMA
        return out;
    };
    if (SCR) {
        // tile 0's sums (its exponents came from the prologue)
#pragma unroll
        for (int st = 0; st < NRW; ++st) row_sum(z, st);
        tile_close();
        for (int t = 0; t + 1 < ntiles; t += 2) {
            iter1s(t, std::integral_constant<int, 1>{});
            if (t + 2 >= ntiles) break;
            iter1s(t + 1, std::integral_constant<int, 0>{});
        }
    } else {
        for (int t = 0; t + 1 < ntiles; t += 2) {
            iter1(t, std::integral_constant<int, 1>{});
            if (t + 2 >= ntiles) break;
            iter1(t + 1, std::integral_constant<int, 0>{});
        }
#pragma unroll
        for (int st = 0; st < NRW; ++st) row_sum(z, st);
        tile_close();
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
        const double s2 = TWO ? ((double)S2 - C2) + ((double)S2o - C2o) : s1;
        // per-point results go to the point's input position
        const long no = (perm && n < N) ? (long)perm[n] : n;
        if (n >= N) {
        } else if (!(s1 > 0.0)) {          // zero column (mpicomm.c:382-390)
            lg = isDA ? -INFINITY : logK;
            lh = logK;
            zc = 1.0;
            pm[no] = -INFINITY;
            logg[no] = lg;
            logh[no] = lh;
        } else {
            const double L1 = log2(s1), L2 = log2(s2);
            const double ma = (double)m + (double)bn;        // the point's max exponent
            lg = kLn2 * (ma + L1);
            lh = kLn2 * (ma * it + L2);
            pm[no] = ma;
            logg[no] = lg;
            logh[no] = lh;
            pbv = (float)(-(double)it * m - L2 + kRLog2);
            ptv = (float)((double)it * ((double)thr - m) - L2 + kRLog2);
        }
        // per 32-point block: 32 exponent offsets pB, then 32 thresholds
        pP[(n / BP) * (2 * BP) + j] = pbv;
        pP[(n / BP) * (2 * BP) + BP + j] = ptv;
    }
    // This wave's 32 points are one pass-2 block: its largest mask threshold
    // tells pass 2 whether the block needs the underflow mask at all.
    float bt = ptv;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bt = fmaxf(bt, __shfl_xor(bt, o));
    if (l == 0) blkThr[bx * (PT / BP) + w] = bt;
    // workgroup partials of sum log g, sum log h and the zero-column count
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        lg += __shfl_xor(lg, o);
        lh += __shfl_xor(lh, o);
        zc += __shfl_xor(zc, o);
    }
    if (l == 0) { wsum[0][w] = lg; wsum[1][w] = lh; wsum[2][w] = zc; }
    __syncthreads();                               // also orders the group maxima stored above
    if (tid < 3)
        partLog[bx * 3 + tid] =
            (wsum[tid][0] + wsum[tid][1]) + (wsum[tid][2] + wsum[tid][3]);
    // Bounds for pass 2: for each centre group g, the largest pass-2 exponent
    // it * z + pB that any (centre of g, point of this block) pair can reach,
    // plus a rounding margin (pass 2 recomputes the exponents in another
    // order).  Pass 2 skips a (group, block) pair whose bound shows every R
    // there rounds to zero.  Lanes of half h take the groups g = h mod 2.
    if (GMX && bnd) {
        // the exchange runs on every lane: a lane reading a lane that is off
        // (as under `h ? shuffle : own`) gets 0, not the value
        const float pbo = __shfl_xor(pbv, 32);
        const float pb = h ? pbo : pbv;
        const int ngr = Kp / CT;
        const long blk = bx * (PT / BP) + wu, nblk = Np / BP;
        const float *gp = gmx + bx * PT + 32 * wu + j;
        constexpr int GB = 8;                      // groups in flight per half-wave
        for (int g0 = 0; g0 < ngr; g0 += 2 * GB) {
            float a[GB];
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                const int g = g0 + 2 * u + h;
                a[u] = (g < ngr) ? gp[(long)g * Np] : -INFINITY;
            }
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                const int g = g0 + 2 * u + h;
                const float ai = it * a[u];
                float v = ai + pb;
                v = (v > -INFINITY) ? v + 1.0f + 1e-5f * (fabsf(ai) + fabsf(pb)) : -INFINITY;
#pragma unroll
                for (int o = 16; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
                if (j == 0 && g < ngr) bnd[(long)g * nblk + blk] = v;
            }
        }
    }
}

template <int KS, int TT, bool TWO>
// (three waves per SIMD where the registers allow: at most 168 each; with the
// screen's branches the D <= 192 instantiation needs 174 and runs two)
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(KS <= 11 ? 3 : 2))) void k_e1_mfma(
    const _Float16 *__restrict__ T16, const float *__restrict__ tt32, long N, long Np,
    const _Float16 *__restrict__ Y16, const float *__restrict__ aK32, int Kp,
    const double *__restrict__ mscal, float beta, float it, int isDA, double logK,
    float *__restrict__ pP, float *__restrict__ blkThr, double *__restrict__ pm,
    double *__restrict__ logg, double *__restrict__ logh, double *__restrict__ partLog,
    const int *__restrict__ perm, float *__restrict__ gmx, float *__restrict__ bnd,
    const int *__restrict__ p1start, float smargin) {
    // the sweep starts at this workgroup's home group (see dg_mixed_order)
    // (kept a scalar: an integer remainder would go through the VALU and turn
    // every DMA address of the sweep into a vector one)
    const int s0 = p1start ? __builtin_amdgcn_readfirstlane(p1start[blockIdx.x]) & ~3 : 0;
    const int t0 = (s0 >= 0 && s0 < Kp / KT) ? s0 : 0;
    // (the screen pays where the cross term takes several MFMAs: D > 48)
    e1_body<KS, TT, TWO, true, TT == 1 && KS >= 4>(blockIdx.x, threadIdx.x, T16, tt32, N, Np, Y16, aK32, Kp, mscal,
                                        beta, it, isDA, logK, pP, blkThr, pm, logg, logh, partLog,
                                        perm, gmx, bnd, t0, smargin);
}

// scal[0..2] = sums over the pass-1 workgroup partials (fixed tree order).
__device__ __forceinline__ void k_sum_partlog_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ part,
    long nb, double *__restrict__ scal) {
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

__global__ __launch_bounds__(1024) void k_sum_partlog(const double *__restrict__ part, long nb,
                                                      double *__restrict__ scal) {
    k_sum_partlog_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, part, nb, scal);
}

// Per-centre exponent term of pass 1, -beta log2(e) |y_k|^2 / 2 (the same
// FP32 expression pass 2 uses), -inf for padding centres, stored per 32-centre
// tile in AKS floats: the tile's 32 values, their largest (pass 1's screen),
// three -inf pads.
__device__ __forceinline__ void k_make_ak_body(
    const int bx, const int by, const int bz, const int gx, const float *__restrict__ yy32,
    int K, int Kp, float beta, float *__restrict__ aK32) {
    const int k = bx * blockDim.x + threadIdx.x;
    const float cb = -0.5f * beta * kLog2e;
    const float v = (k < K) ? cb * yy32[k] : -INFINITY;
    const int tile = k / KT, q = k & (KT - 1);
    float mx = v;                                  // every lane takes part in the exchange
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (k < Kp) {
        aK32[(long)tile * AKS + q] = v;
        if (q < 4) aK32[(long)tile * AKS + KT + q] = (q == 0) ? mx : -INFINITY;
    }
}

__global__ void k_make_ak(const float *__restrict__ yy32, int K, int Kp, float beta,
                          float *__restrict__ aK32) {
    k_make_ak_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, yy32, K, Kp, beta, aK32);
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
// that row.  The work items, (point chunk, centre group) pairs, are mapped so
// that the workgroups the hardware places on one XCD (linear id mod 8) take
// the groups of one chunk after another, keeping that chunk's data in the
// XCD's L2.  k_e2_mfma reads its item from a table (k_item_table) that lists,
// per XCD, only the pairs with blocks to visit: the hardware hands out
// workgroups in order, and pairs with nothing to do interleaved with long
// ones would hold back the dispatch of the next workgroups.  kb_e2_mfma takes
// its items from a work queue over many instances.  The body of one item:
// centre group kg of point chunk cidx.
template <int KS, int TT, bool GM, bool RLO>
__device__ __forceinline__ void e2_body(
    const long cidx, const int kg, const int tid, const _Float16 *__restrict__ T16, const _Float16 *__restrict__ T16t, long Np,
    const _Float16 *__restrict__ Y16, const float *__restrict__ yy32, int K, int Kp, int D,
    const double *__restrict__ mscal, float beta, float it, const float *__restrict__ pP,
    const float *__restrict__ blkThr, long chunk, int KG,
    float *__restrict__ partRT, float *__restrict__ partG, const int2 *__restrict__ brange) {
    constexpr int ld = 16 * KS;
    constexpr int LR = ld + PAD;                   // data row in LDS (halves), padding last
    constexpr int RC = LR / 8;                     // 16-byte chunks per data row
    constexpr int DB = (ld + 31) / 32;
    constexpr int DR = 32 * DB;
    constexpr int TR = BP + PAD;                   // transposed row in LDS (halves)
    constexpr int TC = TR / 8;
    constexpr int NIA = (TT * BP * RC + 63) / 64;  // 1 KB LDS-DMA loads: data rows,
    constexpr int NIB = (TT * DR * TC + 63) / 64;  // transposed rows,
    constexpr int NI = NIA + NIB + 1;              // pB and pThr
    constexpr int NIW = (NI + 3) / 4;              // per wave
    constexpr int SLOT = 4 * NIW * 512;            // halves per ring slot
    constexpr int OFF_TT = NIA * 512, OFF_P = (NIA + NIB) * 512;
    constexpr int NS = 3;                          // ring slots
    constexpr int NX = (TT + 1) * KS;              // cross-term MFMAs per block
    constexpr int RG = TT + (RLO ? 1 : 0);         // R T products per 16 points and d-block
    constexpr int NR = RG * 2 * DB;                // R T MFMAs per block
    static_assert(NS * SLOT * 2 <= 160 * 1024, "LDS ring");
    __shared__ __attribute__((aligned(1024))) _Float16 smem[NS][SLOT];

    // (the & 3 bounds the wave index for the compiler when tid is opaque, as in
    // the persistent kernels: it then still proves each LDS-DMA target disjoint
    // from the LDS reads of the other ring slots and waits for no DMA there)
    const int w = (tid >> 6) & 3, l = tid & 63, h = l >> 5, j = l & 31;
    const int wu = __builtin_amdgcn_readfirstlane(w) & 3;
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
    // The blocks this workgroup visits: the runs of the chunk's blocks that
    // k_block_ranges listed (with the points ordered by their nearest centre,
    // the blocks where some R of this centre group survives in FP16 are few
    // and mostly adjacent; outside the runs every R there rounds to zero, so
    // leaving those blocks out adds nothing).  The accumulators carry over
    // from run to run; each run restarts the staging pipeline.
    // (read through a global address-space view and made wave-uniform, so
    // that the batched kernel, whose pointers come from the instance table,
    // keeps scalar DMA bases; the same for the mask thresholds below)
    typedef const __attribute__((address_space(1))) int *gin;
    typedef const __attribute__((address_space(1))) float *gfl;
    const gin hdr = (gin)brange + 2 * RS * (cidx * KG + kg);
    const int nruns = __builtin_amdgcn_readfirstlane(hdr[0]);
    int rgx = 0, nblocks = 0;                      // this run: first block, count
    long nb = cidx * chunk;                        // its first point
    const _Float16 *T16tc = T16t + cidx * (long)(TT * DR) * chunk;  // T^T of the run

    // Staging: block q goes global -> LDS slot s by LDS-DMA.  The slot is a
    // sequence of 16-byte chunks [data rows | transposed rows | pB, pThr];
    // DMA load I fills chunks 64 I .. 64 I + 63 in lane order, wave w issues
    // loads w, w + 4, ...  Padding chunks (and the tail) take a copy of the
    // load's first source chunk: never read.  Source offsets are fixed per
    // lane; only the three block bases move.
    unsigned voff[NIW];
#pragma unroll
    for (int i = 0; i < NIW; ++i) {
        const int I = wu + 4 * i;
        unsigned o = 0;
        if (I < NIA) {
            const int g = I * 64 + l;
            const int t = g / (BP * RC), rem = g % (BP * RC), row = rem / RC, col = rem % RC;
            if (t < TT && col < ld / 8) o = (unsigned)((((long)t * Np + row) * ld + col * 8) * 2);
        } else if (I < NIA + NIB) {
            const int g = (I - NIA) * 64 + l;
            const int t = g / (DR * TC), rem = g % (DR * TC), row = rem / TC, col = rem % TC;
            if (t < TT && col < BP / 8) o = (unsigned)((((long)t * DR + row) * chunk + col * 8) * 2);
        } else if (I == NIA + NIB) {
            if (l < 2 * BP / 4) o = (unsigned)(l * 16);
        }
        voff[i] = o;
    }
    // per DMA load of this wave: the source base at the run's first block
    // (set per run) and its step per block (both wave-uniform)
    const char *dbase[NIW];
    unsigned dstep[NIW];
#pragma unroll
    for (int i = 0; i < NIW; ++i) {
        const int I = wu + 4 * i;
        dbase[i] = nullptr;
        dstep[i] = (I < NIA) ? (unsigned)(BP * ld * 2) : (I < NIA + NIB) ? (unsigned)(BP * 2)
                                                                       : (unsigned)(2 * BP * 4);
    }
    auto dma = [&](int q, int s, int i) {
        const unsigned qc = (unsigned)(q < nblocks ? q : nblocks - 1);
        const int I = wu + 4 * i;
        __builtin_amdgcn_global_load_lds(
            (const void *)(dbase[i] + qc * dstep[i] + voff[i]),
            (__attribute__((address_space(3))) void *)&smem[s][I * 512], 16, 0, 0);
    };
    // Marks the lane offsets as redefined once per loop iteration (an empty
    // asm, no instruction).  Otherwise the compiler hoists base + offset into
    // a 64-bit VGPR pair per load and adds the block step with a VALU
    // v_lshl_add_u64 before every DMA; with the offsets opaque, the address
    // stays a scalar base (SGPR pair, SALU arithmetic) plus a 32-bit lane
    // offset, the saddr form of global_load_lds.
    auto dma_offsets_opaque = [&]() {
#pragma unroll
        for (int i = 0; i < NIW; ++i) asm volatile("" : "+v"(voff[i]));
    };
    // A operands: data rows (cross term) and transposed rows (R T)
    auto ldT = [&](int s, int t, int ks) {
        return *(const h8 *)&smem[s][(t * BP + j) * LR + 16 * ks + 8 * h];
    };
    auto ldTt = [&](int s, int t, int g) {         // g = 2 db + sh
        return *(const h8 *)&smem[s][OFF_TT + (t * DR + 32 * (g >> 1) + j) * TR + 16 * (g & 1) + 8 * h];
    };
    auto ldP = [&](int s, int which, int m4) {     // which: 0 pB, 1 pThr
        return *(const f4v *)&smem[s][OFF_P + 2 * (which * BP + 8 * m4 + 4 * h)];
    };
    // cross-term MFMA number m (m = (TT + 1) ks + which)
    // HL: every hi-term MFMA first, then the lo terms -- the order of pass 1's
    // screened sweep (exact data, D > 48), so that both passes form
    // bit-identical exponents; otherwise hi and lo alternate per k step, as in
    // pass 1 without the screen
    constexpr bool HL = TT == 1 && KS >= 4;
    auto xks = [&](int m) { return HL ? (m < KS ? m : m - KS) : m / (TT + 1); };
    auto xmfma = [&](int m, const h8 &a0, const h8 &a1, f16v &c) {
        const int ks = xks(m), wq = HL ? (m < KS ? 0 : 1) : m % (TT + 1);
        if (wq == 0) c = mfma(a0, yb[0][ks], c);
        else if (wq == 1) c = mfma(a0, yb[1][ks], c);
        else c = mfma(a1, yb[0][ks], c);
    };
    // exponent (base 2, R * 2^14 scale) of element r from the cross term
    auto expo = [&](const f16v &c, const f4v (&pb)[4], float (&e)[16], int r) {
        e[r] = fmaf(itc1, c[r], aKi) + pb[r >> 2][r & 3];
    };
    // exp-before-pow underflow: elements below their point's threshold are 0.
    // Called under a branch; the empty asm makes the thresholds opaque so the
    // compares cannot be hoisted out of it.
    auto mask = [&](int s, float (&e)[16]) {
#pragma unroll
        for (int m4 = 0; m4 < 4; ++m4) {
            f4v pt = ldP(s, 1, m4);
            asm volatile("" : "+v"(pt));
#pragma unroll
            for (int q = 0; q < 4; ++q)
                e[4 * m4 + q] = (e[4 * m4 + q] < pt[q]) ? -INFINITY : e[4 * m4 + q];
        }
    };
    // R = 2^e as FP16 hi/lo pairs in the B-operand order of the R T product,
    // in 24 steps, three per pair p: the exponentials of elements 2p and
    // 2p+1, then the packs hi = RNE(R) and lo = RNE(R - hi) (v_fma_mix reads
    // one FP16 half of the packed hi and writes one half of lo).  Without
    // RLO, R is the single term hi.
    constexpr int NRS = 24;
    float Rv[16];
    unsigned wh[8], wl[8];
    float gs = 0.f;
    auto resp_step = [&](const float (&e)[16], int st) {
        const int p = st / 3, q = st % 3;
        if (q < 2) {
            Rv[2 * p + q] = ex2(e[2 * p + q]);
            if (!GM) gs += Rv[2 * p + q];
        } else {
            const h2 hp = {(_Float16)Rv[2 * p], (_Float16)Rv[2 * p + 1]};
            const unsigned hb = __builtin_bit_cast(unsigned, hp);
            wh[p] = hb;
            if (RLO) {
                // mixlo writes the low half and mixhi the high half of lb, so
                // lb needs no initial value (early clobber: no input shares it)
                unsigned lb;
                asm("v_fma_mixlo_f16 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=&v"(lb) : "v"(hb), "v"(Rv[2 * p]));
                asm("v_fma_mixhi_f16 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
                    : "+v"(lb) : "v"(hb), "v"(Rv[2 * p + 1]));
                wl[p] = lb;
            }
        }
    };
    auto resp_ops = [&](h8 (&rhi)[2], h8 (&rlo)[2]) {
        // The compiler's hazard recognizer does not see the v_fma_mix writes
        // inside the asm, so it would not insert the two wait states an MFMA
        // needs before reading a VGPR that a VALU instruction has just
        // written: fence the scheduler and wait here, once per block.
        if (RLO) {
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("s_nop 1");
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int sh = 0; sh < 2; ++sh) {
            const u4 a = {wh[4 * sh], wh[4 * sh + 1], wh[4 * sh + 2], wh[4 * sh + 3]};
            rhi[sh] = __builtin_bit_cast(h8, a);
            if (RLO) {
                const u4 b = {wl[4 * sh], wl[4 * sh + 1], wl[4 * sh + 2], wl[4 * sh + 3]};
                rlo[sh] = __builtin_bit_cast(h8, b);
            } else {
                rlo[sh] = rhi[sh];                 // unused
            }
        }
        if (!GM) { kahan(G, Gc, gs); gs = 0.f; }
    };
    // R T MFMAs: accumulator d-block db, 16-point half sh, term wq (hi R,
    // lo R when RLO, then hi R with the second data term).  Issue order: by sh, then by pairs of d-blocks,
    // alternating the two accumulators of a pair so that no MFMA waits for
    // the one just before it.
    struct RtOp { int db, sh, wq, pairFirst; };
    auto rt_op = [&](int i) {
        constexpr int GP = RG;
        const int sh = i / (DB * GP), r = i % (DB * GP);
        const int pair = r / (2 * GP), in = r % (2 * GP);
        const int npd = (2 * pair + 1 < DB) ? 2 : 1;
        return RtOp{2 * pair + in % npd, sh, in / npd, in == 0};
    };
    auto rmfma = [&](const RtOp &o, const h8 &u0, const h8 &u1, const h8 (&rhi)[2],
                     const h8 (&rlo)[2]) {
        if (o.wq == 0) acc[o.db] = mfma(u0, rhi[o.sh], acc[o.db]);
        else if (RLO && o.wq == 1) acc[o.db] = mfma(u0, rlo[o.sh], acc[o.db]);
        else acc[o.db] = mfma(u1, rhi[o.sh], acc[o.db]);
    };
    // whole-block versions (prologue and epilogue)
    auto cross_exp = [&](int s, float (&e)[16]) {
        f16v c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.f;
#pragma unroll
        for (int m = 0; m < NX; ++m) {
            const int ks = xks(m);
            xmfma(m, ldT(s, 0, ks), ldT(s, TT - 1, ks), c);
        }
        f4v pb[4];
#pragma unroll
        for (int m4 = 0; m4 < 4; ++m4) pb[m4] = ldP(s, 0, m4);
#pragma unroll
        for (int r = 0; r < 16; ++r) expo(c, pb, e, r);
    };
    auto resp_rt = [&](int s, const float (&e)[16]) {
#pragma unroll
        for (int st = 0; st < NRS; ++st) resp_step(e, st);
        h8 rhi[2], rlo[2];
        resp_ops(rhi, rlo);
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const RtOp o = rt_op(i);
            const int g = 2 * o.db + o.sh;
            rmfma(o, ldTt(s, 0, g), ldTt(s, TT - 1, g), rhi, rlo);
        }
    };
    long blk0 = 0;                                 // index of the run's first block
    const gfl thrg = (gfl)blkThr;

    for (int run = 0; run < nruns; ++run) {
        rgx = __builtin_amdgcn_readfirstlane(hdr[2 + 2 * run]);
        nblocks = __builtin_amdgcn_readfirstlane(hdr[3 + 2 * run]);
        nb = cidx * chunk + (long)rgx * BP;
        T16tc = T16t + cidx * (long)(TT * DR) * chunk + (long)rgx * BP;
        blk0 = nb / BP;
#pragma unroll
        for (int i = 0; i < NIW; ++i) {
            const int I = wu + 4 * i;
            dbase[i] = (I < NIA) ? (const char *)(T16 + nb * ld)
                     : (I < NIA + NIB) ? (const char *)T16tc
                                       : (const char *)(pP + blk0 * (2 * BP));
        }
        // the previous run's last reads of the ring come before this run's DMA
        if (run > 0) __syncthreads();
#pragma unroll
        for (int i = 0; i < NIW; ++i) dma(0, 0, i);
#pragma unroll
        for (int i = 0; i < NIW; ++i) dma(1, 1, i);
        __syncthreads();
        float e[16];
        cross_exp(0, e);
        if (thrg[blk0] > -kMaskMargin) mask(0, e);
        // Steady state, one block per iteration, in a fixed issue order
        // (sched_barrier keeps the compiler from regrouping it):
        //   phase 1: cross-term MFMA m of block b+1, then step m of: block b's
        //            exponentials and packs (16 steps) and the DMA loads of
        //            block b+2;
        //   phase 2: R T MFMA i of block b, then step i of block b+1's
        //            exponents (pB loads, 16 elements).
        // The A operands are read from LDS two MFMAs ahead.
        // work steps are spread evenly over the MFMA slots: slot m of a phase
        // with W steps runs steps m W / n .. (m + 1) W / n - 1
        constexpr int W1 = NRS + NIW;              // phase 1: R steps, DMA loads
        constexpr int W2 = 4 + 16;                 // phase 2: pB loads, exponents
        // One iteration per block b; the ring slots are template constants, so
        // LDS addresses fold into the instructions' offset fields (the loop
        // below runs three iterations per trip, one per slot rotation).
        auto iter = [&](const int b, auto S0) {
            constexpr int s0 = decltype(S0)::value, s1 = (s0 + 1) % NS, s2 = (s0 + 2) % NS;
            dma_offsets_opaque();
            f16v c;
#pragma unroll
            for (int r = 0; r < 16; ++r) c[r] = 0.f;
            h8 xa0[KS], xa1[KS];
#pragma unroll
            for (int ks = 0; ks < 2 && ks < KS; ++ks) {
                xa0[ks] = ldT(s1, 0, ks);
                if (TT == 2) xa1[ks] = ldT(s1, 1, ks);
            }
            // R T A operands (block b's transposed rows, resident all iteration),
            // indexed g = 2 db + sh: the first pair's are read three MFMAs
            // before phase 1 ends (at its first MFMA when phase 1 has fewer
            // than three), each pair's first MFMA reads the next pair's
            constexpr int MLD = NX >= 3 ? NX - 3 : 0;
            h8 ru0[2 * DB], ru1[2 * DB];
            auto ld_pair = [&](int i0) {           // operands of the pair starting at MFMA i0
                const RtOp o = rt_op(i0);
                const int npd = (o.db + 1 < DB && rt_op(i0 + 1).db != o.db) ? 2 : 1;
#pragma unroll
                for (int x = 0; x < 2; ++x)
                    if (x < npd) {
                        const int g = 2 * (o.db + x) + o.sh;
                        ru0[g] = ldTt(s0, 0, g);
                        if (TT == 2) ru1[g] = ldTt(s0, 1, g);
                    }
            };
#pragma unroll
            for (int m = 0; m < NX; ++m) {
                const int ks = xks(m);
                if (HL) {                          // each half reads the data rows again
                    if (m + 2 < NX) xa0[xks(m + 2)] = ldT(s1, 0, xks(m + 2));
                } else if (m % (TT + 1) == 0 && ks + 2 < KS) {
                    xa0[ks + 2] = ldT(s1, 0, ks + 2);
                    if (TT == 2) xa1[ks + 2] = ldT(s1, 1, ks + 2);
                }
                xmfma(m, xa0[ks], TT == 2 ? xa1[ks] : xa0[ks], c);
#pragma unroll
                for (int st = m * W1 / NX; st < (m + 1) * W1 / NX; ++st) {
                    if (st < NRS) resp_step(e, st);
                    else dma(b + 2, s2, st - NRS);
                }
                if (m == MLD) ld_pair(0);
                __builtin_amdgcn_sched_barrier(0);
            }
            h8 rhi[2], rlo[2];
            resp_ops(rhi, rlo);
            f4v pb[4];
            float en[16];
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                const RtOp o = rt_op(i);
                if (o.pairFirst) {
                    // the next pair starts after this one's 2 RG MFMAs (or fewer)
                    int nx = i + 1;
                    while (nx < NR && !rt_op(nx).pairFirst) ++nx;
                    if (nx < NR) ld_pair(nx);
                }
                const int g = 2 * o.db + o.sh;
                rmfma(o, ru0[g], TT == 2 ? ru1[g] : ru0[g], rhi, rlo);
#pragma unroll
                for (int st = i * W2 / NR; st < (i + 1) * W2 / NR; ++st) {
                    if (st < 4) pb[st] = ldP(s1, 0, st);
                    else expo(c, pb, en, st - 4);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            // rarely taken; a real branch after every MFMA of the iteration.  The
            // index goes through readfirstlane so that the load stays a scalar
            // one: in the unrolled loop the compiler otherwise makes it a vector
            // load whose vmcnt(0) wait exposes a memory round trip per block.
            if (thrg[blk0 + __builtin_amdgcn_readfirstlane(b + 1)] > -kMaskMargin) mask(s1, en);
#pragma unroll
            for (int r = 0; r < 16; ++r) e[r] = en[r];
            __syncthreads();                       // also waits for this iteration's DMA
        };
        static_assert(NS == 3, "the loop unrolls one trip per slot rotation");
        for (int b = 0; b + 1 < nblocks; b += 3) {
            iter(b, std::integral_constant<int, 0>{});
            if (b + 2 >= nblocks) break;
            iter(b + 1, std::integral_constant<int, 1>{});
            if (b + 3 >= nblocks) break;
            iter(b + 2, std::integral_constant<int, 2>{});
        }
        resp_rt((nblocks - 1) % NS, e);
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

template <int KS, int TT, bool GM, bool RLO>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_e2_mfma(
    const _Float16 *__restrict__ T16, const _Float16 *__restrict__ T16t, long Np,
    const _Float16 *__restrict__ Y16, const float *__restrict__ yy32, int K, int Kp, int D,
    const double *__restrict__ mscal, float beta, float it, const float *__restrict__ pP,
    const float *__restrict__ blkThr, long chunk, int KG,
    float *__restrict__ partRT, float *__restrict__ partG, const int2 *__restrict__ brange,
    const int *__restrict__ itab, const int *__restrict__ icnt) {
    const int x = blockIdx.x & 7, i = blockIdx.x >> 3;
    if (i >= icnt[x]) return;                      // this XCD's list is shorter
    const int lid = itab[blockIdx.x];
    e2_body<KS, TT, GM, RLO>(lid / KG, lid % KG, threadIdx.x, T16, T16t, Np, Y16, yy32, K, Kp, D,
                             mscal, beta, it, pP, blkThr, chunk, KG, partRT, partG, brange);
}

// The blocks pass 2 visits for one (point chunk, centre group) pair, lid =
// chunk * KG + group: the blocks whose pass-1 bound is not below `below`
// (below it, every R of the group there rounds to zero) as runs of adjacent
// blocks, dead gaps of up to RGAP blocks merged into a run and the last run
// taking in everything after it once MAXR runs are used; the whole chunk as
// one run when bnd is null.  brange[RS lid] = {runs, blocks}, then the runs
// {first block within the chunk, count}.  One workgroup per pair: the waves
// ballot a segment of NTH blocks, the first lane walks the live ones.
__device__ __forceinline__ void block_range_body(const int lid, const float *__restrict__ bnd,
                                                 long Np, long chunk, int KG, float below,
                                                 int2 *__restrict__ brange) {
    __shared__ unsigned long long s_live[NTH / 64];
    const long cidx = lid / KG;
    const int kg = lid % KG;
    const long nb = cidx * chunk;
    const long ne = (nb + chunk < Np) ? nb + chunk : Np;
    const int nblocks = (ne > nb) ? (int)((ne - nb) / BP) : 0;
    int2 *out = brange + (long)RS * lid;
    if (!bnd) {
        if (threadIdx.x == 0) {
            out[0] = make_int2(nblocks > 0 ? 1 : 0, nblocks);
            out[1] = make_int2(0, nblocks);
        }
        return;
    }
    const float *bb = bnd + (long)kg * (Np / BP) + nb / BP;
    const int w = threadIdx.x >> 6;
    int nr = 0, cs = -1, cl = -1, tot = 0;         // first lane: runs, current run, blocks
    for (int q0 = 0; q0 < nblocks; q0 += NTH) {
        const int q = q0 + (int)threadIdx.x;
        const unsigned long long bal = __ballot(q < nblocks && !(bb[q] < below));
        if ((threadIdx.x & 63) == 0) s_live[w] = bal;
        __syncthreads();
        if (threadIdx.x == 0) {
            // one stretch of adjacent live blocks at a time: from its first
            // block s, the stretch runs to the first dead block after it
            for (int x = 0; x < NTH / 64; ++x) {
                unsigned long long m = s_live[x];
                while (m) {
                    const int s0 = __builtin_ctzll(m);
                    const unsigned long long dead = ~m >> s0;
                    const int len = dead ? __builtin_ctzll(dead) : 64 - s0;
                    const int b = q0 + 64 * x + s0;
                    m = (s0 + len < 64) ? m & (~0ull << (s0 + len)) : 0ull;
                    if (cs < 0) {
                        cs = b;
                    } else if (b - cl - 1 > RGAP && nr < MAXR - 1) {
                        out[1 + nr] = make_int2(cs, cl - cs + 1);
                        tot += cl - cs + 1;
                        ++nr;
                        cs = b;
                    }
                    cl = b + len - 1;
                }
            }
        }
        __syncthreads();                           // s_live is rewritten next segment
    }
    if (threadIdx.x == 0) {
        if (cs >= 0) {
            out[1 + nr] = make_int2(cs, cl - cs + 1);
            tot += cl - cs + 1;
            ++nr;
        }
        out[0] = make_int2(nr, tot);
    }
}

__global__ __launch_bounds__(NTH) void k_block_ranges(const float *__restrict__ bnd, long Np,
                                                       long chunk, int KG, float below,
                                                       int2 *__restrict__ brange) {
    block_range_body(blockIdx.x, bnd, Np, chunk, KG, below, brange);
}

// Pass 2's work list, one workgroup per XCD x: the (chunk, group) pairs of
// chunks c = x, x + 8, ... (chunk after chunk, groups in order) whose block
// range is not empty, at itab[x + 8 i] (i = 0, 1, ...), and their number
// icnt[x].  Ballots and per-wave counts keep the order.
__global__ __launch_bounds__(NTH) void k_item_table(const int2 *__restrict__ brange, int NC,
                                                     int KG, int *__restrict__ itab,
                                                     int *__restrict__ icnt) {
    __shared__ int s_w[NTH / 64];
    const int x = blockIdx.x;
    const int nch = (NC - x + 7) / 8;              // this XCD's chunks
    const int total = nch * KG;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    int base = 0;
    for (int q0 = 0; q0 < total; q0 += NTH) {
        const int q = q0 + (int)threadIdx.x;
        const int lid = (q / KG * 8 + x) * KG + q % KG;
        const bool live = q < total && brange[(long)RS * lid].y > 0;
        const unsigned long long bal = __ballot(live);
        const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
        if (l == 0) s_w[w] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int y = 0; y < w; ++y) off += s_w[y];
        if (live) itab[x + 8 * (off + pos)] = lid;
        for (int y = 0; y < NTH / 64; ++y) base += s_w[y];
        __syncthreads();
    }
    if (threadIdx.x == 0) icnt[x] = base;
}

// ---------------------------------------------------------------- helpers
// RT[k][d] = sum_c partRT[c][d][k] and G[k] = sum_c partG[c][k], each in FP64
// in chunk order.  One thread per output (lanes along the contiguous centre
// axis); the chunk loop issues eight independent loads before it adds them,
// in order.  A workgroup owns 64 centres x 16 dims and transposes them through
// LDS so each FP64 row segment of RT is written as 128 contiguous bytes.
// Workgroups of the first dim tile also sum G.
constexpr int RK = 64, RD = 16, RU = 8;

__device__ __forceinline__ void k_reduce32_body(
    const int bx, const int by, const int bz, const int gx, const float *__restrict__ partRT,
    const float *__restrict__ partG, int NC, int K, int D, double *__restrict__ RT,
    double *__restrict__ G, const int2 *__restrict__ brange) {
    __shared__ double tile[RK][RD + 1];
    const int t = threadIdx.x, kk = t & (RK - 1), dd = t / RK;
    const int k0 = bx * RK, d0 = by * RD;
    const int k = k0 + kk, d = d0 + dd;
    const long KD = (long)K * D;
    // chunk c's partial sums of this centre group exist only where pass 2 had
    // blocks to visit (blocks > 0 in the pair's run-table entry; no table:
    // everywhere)
    const int KG = (K + CT - 1) / CT, grp = k0 / CT;
    auto live = [&](int c) { return !brange || brange[RS * ((long)c * KG + grp)].y > 0; };
    double s = 0.0;
    if (k < K && d < D) {
        const float *p = partRT + (long)d * K + k;
        int c = 0;
        for (; c + RU <= NC; c += RU) {
            float v[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) v[u] = p[(long)(c + u) * KD];
#pragma unroll
            for (int u = 0; u < RU; ++u) s += live(c + u) ? (double)v[u] : 0.0;
        }
        for (; c < NC; ++c) s += live(c) ? (double)p[(long)c * KD] : 0.0;
    }
    tile[kk][dd] = s;
    __syncthreads();
    {
        const int r = t / RD, e = t % RD;
        if (k0 + r < K && d0 + e < D) RT[(long)(k0 + r) * D + d0 + e] = tile[r][e];
    }
    if (by == 0 && t < RK && k < K) {
        // RU loads in flight, then their sums in chunk order (as for R T)
        double g = 0.0;
        int c = 0;
        for (; c + RU <= NC; c += RU) {
            float v[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) v[u] = partG[(long)(c + u) * K + k];
#pragma unroll
            for (int u = 0; u < RU; ++u) g += live(c + u) ? (double)v[u] : 0.0;
        }
        for (; c < NC; ++c) g += live(c) ? (double)partG[(long)c * K + k] : 0.0;
        G[k] = g;
    }
}

__global__ __launch_bounds__(RK * RD) void k_reduce32(const float *__restrict__ partRT,
                                                      const float *__restrict__ partG, int NC,
                                                      int K, int D, double *__restrict__ RT,
                                                      double *__restrict__ G,
                                                      const int2 *__restrict__ brange) {
    k_reduce32_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, partRT, partG, NC, K, D, RT, G,
                    brange);
}

// Zero-column fall-back R = 1/K for the points pass 2 skipped: RT[k] +=
// (1/K) sum_{n in Z} t_n, G[k] += |Z| / K.  Exits at once when |Z| = 0.
__device__ __forceinline__ void k_zero_fix_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ scal,
    const double *__restrict__ T64, int ldt, const double *__restrict__ pm, long N, int K,
    int D, double *__restrict__ RT, double *__restrict__ G) {
    const double nz = scal[kScalNZero];
    if (nz == 0.0) return;
    __shared__ double red[256];
    const int d = bx;
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

__global__ __launch_bounds__(256) void k_zero_fix(const double *__restrict__ scal,
                                                  const double *__restrict__ T64, int ldt,
                                                  const double *__restrict__ pm, long N, int K,
                                                  int D, double *__restrict__ RT,
                                                  double *__restrict__ G) {
    k_zero_fix_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, scal, T64, ldt, pm, N, K, D, RT,
                    G);
}

// Workgroup partials of max |x| and of the count of values not exact in
// FP16 over rows x D (stride ld): part[2 b], part[2 b + 1].
constexpr int kAmaxWG = 256;
__device__ __forceinline__ void k_absmax_part_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ X,
    long rows, int ld, int D, double *__restrict__ part) {
    __shared__ double smax[256], scnt[256];
    double mx = 0.0, cnt = 0.0;
    const long total = rows * D;
    for (long e = (long)bx * 256 + threadIdx.x; e < total; e += (long)gx * 256) {
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
    if (threadIdx.x == 0) { part[2 * bx] = smax[0]; part[2 * bx + 1] = scnt[0]; }
}

__global__ __launch_bounds__(256) void k_absmax_part(const double *__restrict__ X, long rows,
                                                     int ld, int D, double *__restrict__ part) {
    k_absmax_part_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, X, rows, ld, D, part);
}

__device__ __forceinline__ void k_absmax_final_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ part,
    int n, double *__restrict__ out) {
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

__global__ __launch_bounds__(256) void k_absmax_final(const double *__restrict__ part, int n,
                                                      double *__restrict__ out) {
    k_absmax_final_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, part, n, out);
}

// Power-of-two scale putting max |x| at or below 2^14.
__device__ __forceinline__ double pow2_scale(double amax) {
    if (!(amax > 0.0)) return 1.0;
    const int e = ilogb(amax);                  // 2^e <= amax < 2^(e+1)
    return ldexp(1.0, 13 - e);
}

// Split T (N x D, stride ldt) into FP16 terms: T16[t][n][d], d < ld16, with
// rows n >= N and columns d >= D zero; tt32[n] = |t'|^2 of the split value.
// Row n holds input row perm[n] (row n when perm is null).
__global__ void k_make_t16(const double *__restrict__ T64, int ldt, long N, long Np, int D,
                           int ld16, int TT, double scale, const int *__restrict__ perm,
                           _Float16 *__restrict__ T16, float *__restrict__ tt32) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Np) return;
    const long src = (n < N && perm) ? (long)perm[n] : n;      // input row of point n
    double ss = 0.0;
    for (int d = 0; d < ld16; ++d) {
        const double v = (n < N && d < D) ? T64[src * ldt + d] * scale : 0.0;
        const _Float16 hi = (_Float16)(float)v;
        const _Float16 lo = (TT == 2) ? (_Float16)(float)(v - (double)(float)hi) : (_Float16)0.0f;
        T16[n * ld16 + d] = hi;
        if (TT == 2) T16[(Np + n) * ld16 + d] = lo;
        const double rep = ((double)(float)hi + (double)(float)lo) / scale;
        ss = fma(rep, rep, ss);
    }
    tt32[n] = (float)ss;
}

// Transposed data, chunk-major: for point n in chunk c = n / CL (offset
// o = n % CL), T16t[((c TT + t) DR + d) CL + pos(o)] = T16[t][n][d] for d < DR
// (rows >= ld16 zero), with the points of each 16-point group stored in the
// order 0-3, 8-11, 4-7, 12-15 (pos swaps the two middle quads) so that a
// lane's eight R-operand points are one 16-byte run.  With ones_row >= 0,
// row ones_row of term 0 is 1 for real points (n < N) and 0 for padding.
// 64 x 64 tiles; CL is a multiple of 64.
__global__ void k_transpose16(const _Float16 *__restrict__ T16, long N, long Np, int ld16, int DR,
                              int ones_row, long CL, int TT, _Float16 *__restrict__ T16t) {
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
    const long ch = n0 / CL, o0 = n0 % CL;
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        const int r = e / 64, c = e % 64;           // r: dim, c: stored position
        const int q = (c >> 2) & 3;
        const int src = (c & ~15) | ((((q & 1) << 1) | (q >> 1)) << 2) | (c & 3);
        const int d = d0 + r;
        if (d < DR && n0 + c < Np)
            T16t[((ch * TT + t) * DR + d) * CL + o0 + c] = tile[src][r];
    }
}

// Split Y (K x D, stride ldt, FP64) into two FP16 terms with a power-of-two
// scale chosen from scal_in[0] = max |Y|: Y16[t][k][d], rows k >= K zero;
// yy32[k] = |y'|^2 of the split value; mscal[0] = the scale.
__device__ __forceinline__ void k_make_y16_body(
    const int bx, const int by, const int bz, const int gx, const double *__restrict__ Y64,
    int ldt, int K, int Kp, int D, int ld16, const double *__restrict__ amax,
    _Float16 *__restrict__ Y16, float *__restrict__ yy32, double *__restrict__ mscal) {
    // one wave per centre row; lanes step over d, the norm is a wave sum
    const int k = bx * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    const double scale = pow2_scale(amax[0]);
    if (bx == 0 && threadIdx.x == 0) mscal[0] = scale;
    if (k >= Kp) return;
    double ss = 0.0;
    for (int d = l; d < ld16; d += 64) {
        const double v = (k < K && d < D) ? Y64[(long)k * ldt + d] * scale : 0.0;
        const _Float16 hi = (_Float16)(float)v;
        const _Float16 lo = (_Float16)(float)(v - (double)(float)hi);
        Y16[(long)k * ld16 + d] = hi;
        Y16[((long)Kp + k) * ld16 + d] = lo;
        const double rep = ((double)(float)hi + (double)(float)lo) / scale;
        ss = fma(rep, rep, ss);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    if (l == 0) yy32[k] = (float)ss;
}

__global__ __launch_bounds__(256) void k_make_y16(const double *__restrict__ Y64, int ldt, int K,
                                                  int Kp, int D, int ld16,
                                                  const double *__restrict__ amax,
                                                  _Float16 *__restrict__ Y16,
                                                  float *__restrict__ yy32,
                                                  double *__restrict__ mscal) {
    k_make_y16_body(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, Y64, ldt, K, Kp, D, ld16, amax,
                    Y16, yy32, mscal);
}

// Nearest centre of each point from the first FP16 terms (an ordering key,
// not a result): key[n] = argmax_k (t'_n . y'_k - s |y'_k|^2 / 2), s the
// product of the two power-of-two operand scales (mscal[0] for Y, mscal[1]
// for T), first k on ties.  One thread per point with its row in registers as
// FP16 pairs (v_dot2_f32_f16); the centres pass through LDS 32 at a time.
template <int KS>
__global__ __launch_bounds__(256) void k_nearest(const _Float16 *__restrict__ T16, long N,
                                                 const _Float16 *__restrict__ Y16,
                                                 const float *__restrict__ yy32, int K,
                                                 const double *__restrict__ mscal,
                                                 int *__restrict__ key) {
    constexpr int ld = 16 * KS, NPR = ld / 2, CTL = 32;
    __shared__ h2 ys[CTL][NPR];
    __shared__ float ysq[CTL];
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    h2 t[NPR];
#pragma unroll
    for (int p = 0; p < NPR; ++p) {
        const h2 z = {(_Float16)0.0f, (_Float16)0.0f};
        t[p] = (n < N) ? *(const h2 *)(T16 + n * ld + 2 * p) : z;
    }
    const float s = (float)(0.5 * mscal[0] * mscal[1]);
    float best = -INFINITY;
    int bk = 0;
    for (int k0 = 0; k0 < K; k0 += CTL) {
        __syncthreads();
        for (int e = threadIdx.x; e < CTL * NPR; e += 256) {
            const int r = e / NPR, p = e % NPR;
            const h2 z = {(_Float16)0.0f, (_Float16)0.0f};
            ys[r][p] = (k0 + r < K) ? *(const h2 *)(Y16 + (long)(k0 + r) * ld + 2 * p) : z;
        }
        if (threadIdx.x < CTL) ysq[threadIdx.x] = (k0 + (int)threadIdx.x < K) ? s * yy32[k0 + threadIdx.x] : INFINITY;
        __syncthreads();
        for (int r = 0; r < CTL; ++r) {
            float acc = 0.f;
#pragma unroll
            for (int p = 0; p < NPR; ++p) acc = __builtin_amdgcn_fdot2(t[p], ys[r][p], acc, false);
            const float v = acc - ysq[r];
            if (v > best) { best = v; bk = k0 + r; }
        }
    }
    if (n < N) key[n] = bk;
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
    float *aK32 = nullptr;        // Kp/KT x AKS: pass-1 per-centre exponent terms per tile
    double *amaxPart = nullptr;   // 2 x kAmaxWG
    float *gmx = nullptr;         // Kp/CT x Np: pass-1 maximum exponent per centre group and point
    float *bnd = nullptr;         // Kp/CT x Np/BP: pass-2 exponent bound per group and block
    int2 *brange = nullptr;       // NC x Kp/CT x RS: pass-2 block runs per (chunk, group)
    bool screen = true;           // pass 1 screens tiles (data exact in FP16)
    int *itab = nullptr;          // 8 x Kp/CT x ceil(NC/8): pass-2 work list per XCD
    int *icnt = nullptr;          // 8: its lengths
    long Np = 0;
    long CL = 0;                  // points per pass-2 chunk (multiple of 64)
    int Kp = 0, KS = 0, TT = 1, NC = 1;
};

// Pass 2 skips a (centre group, block) pair whose exponent bound is below
// this: every R there is below 2^-25 after the 2^14 carry, so both FP16
// terms are zero; when G is summed from the FP32 values (no spare d row),
// below 2^-160, so those are zero in FP32 as well.
inline float skip_below(bool gm) { return gm ? -25.0f : -160.0f; }

// DG_MIXED_SKIP=0 makes pass 2 visit every block; DG_MIXED_ORDER=0 keeps the
// points in input order
bool env_off(const char *name) {
    const char *v = std::getenv(name);
    return v && v[0] == '0' && v[1] == 0;
}

MixedBufs *mb(dg_ctx *c) { return reinterpret_cast<MixedBufs *>(c->mix); }

template <int KS, int TT, bool GM>
void launch_e2(dg_ctx *c, MixedBufs *b, float beta, float it, long chunk, int KG) {
    const unsigned grid = (unsigned)(8 * KG * ((b->NC + 7) / 8));
    if (c->cfg.precision == DG_MIXED_R1)
        hipLaunchKernelGGL((k_e2_mfma<KS, TT, GM, false>), dim3(grid), dim3(NTH), 0, c->stream,
                           (const _Float16 *)c->T16, b->T16t, b->Np, b->Y16, c->yy32, c->K, b->Kp,
                           c->D, b->mscal, beta, it, c->pP32, b->blkThr, chunk, KG, c->partRT32,
                           c->partG32, b->brange, b->itab, b->icnt);
    else
        hipLaunchKernelGGL((k_e2_mfma<KS, TT, GM, true>), dim3(grid), dim3(NTH), 0, c->stream,
                           (const _Float16 *)c->T16, b->T16t, b->Np, b->Y16, c->yy32, c->K, b->Kp,
                           c->D, b->mscal, beta, it, c->pP32, b->blkThr, chunk, KG, c->partRT32,
                           c->partG32, b->brange, b->itab, b->icnt);
}

template <int KS, int TT>
void launch_mixed(dg_ctx *c, MixedBufs *b, float beta, float it, int two, int isDA,
                  double logK) {
    // pass 1 leaves out tiles whose every element is below 2^-48 of its
    // point's sums (exponents 48 / it below the running maximum)
    const float smargin = b->screen ? 48.0f / it : INFINITY;
    hipLaunchKernelGGL(k_make_ak, dim3((unsigned)((b->Kp + 255) / 256)), dim3(256), 0, c->stream,
                       c->yy32, c->K, b->Kp, beta, b->aK32);
    if (two)
        hipLaunchKernelGGL((k_e1_mfma<KS, TT, true>), dim3((unsigned)(b->Np / PT)), dim3(NTH), 0,
                           c->stream, (const _Float16 *)c->T16, c->tt32, c->N, b->Np, b->Y16,
                           b->aK32, b->Kp, b->mscal, beta, it, isDA, logK, c->pP32, b->blkThr,
                           c->pm, c->logg, c->logh, b->partLog, c->perm, b->gmx, b->bnd,
                           c->p1start, smargin);
    else
        hipLaunchKernelGGL((k_e1_mfma<KS, TT, false>), dim3((unsigned)(b->Np / PT)), dim3(NTH), 0,
                           c->stream, (const _Float16 *)c->T16, c->tt32, c->N, b->Np, b->Y16,
                           b->aK32, b->Kp, b->mscal, beta, it, isDA, logK, c->pP32, b->blkThr,
                           c->pm, c->logg, c->logh, b->partLog, c->perm, b->gmx, b->bnd,
                           c->p1start, smargin);
    if (c->timing) (void)hipEventRecord(c->ev[1], c->stream);
    const long chunk = b->CL;
    const int KG = b->Kp / CT;
    // G comes out of the R T product when the padded d range has a spare row
    constexpr int DR = 32 * ((16 * KS + 31) / 32);
    hipLaunchKernelGGL(k_block_ranges, dim3((unsigned)(KG * b->NC)), dim3(NTH), 0, c->stream,
                       b->bnd, b->Np, chunk, KG, skip_below(DR > c->D), b->brange);
    hipLaunchKernelGGL(k_item_table, dim3(8), dim3(NTH), 0, c->stream, b->brange, b->NC, KG,
                       b->itab, b->icnt);
    if (DR > c->D) launch_e2<KS, TT, true>(c, b, beta, it, chunk, KG);
    else launch_e2<KS, TT, false>(c, b, beta, it, chunk, KG);
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

// ---------------------------------------------------------------- batched
// The E-step of many instances per step.  Both E-passes run as persistent
// grids (as many workgroups as fit on the device) over a work queue: a
// workgroup takes the next item from a device counter, finds its instance in
// the launch's list (items are numbered instance after instance, each
// instance's first item in its table entry), runs the single-instance
// workgroup body on it, and stops when the queue is empty.  Every item
// writes its own outputs, so results do not depend on the order items run
// in: each instance gets exactly its single-instance results.  The light
// kernels take one grid dimension as the list slot and return at once in
// workgroups past the instance's own grid.

// the list entry whose first item is the largest one <= item (list ordered
// by first item); item is wave-uniform, so this runs on scalar loads
template <bool E2>
__device__ __forceinline__ const BInst &find_inst(const BInst *__restrict__ tab,
                                                  const int *__restrict__ list, int nlist,
                                                  int item) {
    int lo = 0, hi = nlist - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int first = E2 ? tab[list[mid]].item_e2 : tab[list[mid]].item_e1;
        if (first <= item) lo = mid;
        else hi = mid - 1;
    }
    return tab[list[lo]];
}

// next work item for the whole workgroup (uniform), or >= total when none
__device__ __forceinline__ int next_item(int *__restrict__ counter, int *s_item) {
    __syncthreads();                      // every wave is done with the previous item
    if (threadIdx.x == 0) *s_item = atomicAdd(counter, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_item);
}

// the thread index, redefined (an empty asm) in every work-loop iteration:
// the bodies' lane-derived values are then recomputed per item instead of
// being hoisted out of the loop and held in registers across it
__device__ __forceinline__ int opaque_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

template <int KS, int TT, bool TWO>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(2))) void kb_e1_mfma(
    const BInst *__restrict__ tab, const int *__restrict__ list, int nlist, int total,
    int *__restrict__ counter) {
    __shared__ int s_item;
    for (;;) {
        const int item = next_item(counter, &s_item);
        if (item >= total) break;
        const BInst &I = find_inst<false>(tab, list, nlist, item);
        if (!I.do_e) continue;
        // no register room for the group maxima in the batched KS = 8, split-data
        // pass 1 (those instances get no bounds, and pass 2 visits every block);
        // the screen and the start tile as in k_e1_mfma, so that each instance
        // takes exactly its single-instance steps
        const long bx = item - I.item_e1;
        const int s0 = I.p1start ? __builtin_amdgcn_readfirstlane(I.p1start[bx]) & ~3 : 0;
        const int t0 = (s0 >= 0 && s0 < I.Kp / KT) ? s0 : 0;
        const float smargin = I.screen ? 48.0f / I.it : INFINITY;
        e1_body<KS, TT, TWO, !(KS == 8 && TT == 2), TT == 1 && KS >= 4>(
            bx, opaque_tid(), I.T16, I.tt32, I.N, I.Np, I.Y16, I.aK32, I.Kp, I.mscal, I.beta_f, I.it,
            I.isDA, I.logK, I.pP32, I.blkThr, I.pm, I.logg, I.logh, I.partLog, I.perm, I.gmx, I.bnd,
            t0, smargin);
    }
}

template <int KS, int TT, bool GM, bool RLO>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(1, 1))) void kb_e2_mfma(
    const BInst *__restrict__ tab, const int *__restrict__ list, int nlist, int total,
    int *__restrict__ counter) {
    __shared__ int s_item;
    for (;;) {
        const int item = next_item(counter, &s_item);
        if (item >= total) break;
        const BInst &I = find_inst<true>(tab, list, nlist, item);
        if (!I.do_e) continue;
        // the instance's items in XCD-aware order, empty ones included (they
        // write zero partial sums; the queue hands items out dynamically)
        const int L = item - I.item_e2, KGi = I.Kp / CT, xs = L & 7, sl = L >> 3;
        e2_body<KS, TT, GM, RLO>((long)(sl / KGi) * 8 + xs, sl % KGi, opaque_tid(), I.T16, I.T16t, I.Np, I.Y16, I.yy32, I.K, I.Kp,
                                 I.D, I.mscal, I.beta_f, I.it, I.pP32, I.blkThr, I.CL,
                                 I.Kp / CT, I.partRT32, I.partG32, I.brange);
    }
}

__global__ void kb_zero_ints(int *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static_assert(RK == kBRK && RD == kBRD && kAmaxWG == kBAmaxWG && CT == 128,
              "flat-grid tile sizes");

// Flat-grid kernels: workgroup blockIdx.x belongs to instance map[.].x and is
// its workgroup map[.].y (grids sized by bk_count for each instance).
__global__ __launch_bounds__(256) void kb_make_ak(const BInst *__restrict__ tab,
                                                  const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_e) return;
    k_make_ak_body(m.y, 0, 0, 0, I.yy32, I.K, I.Kp, I.beta_f, I.aK32);
}

__global__ __launch_bounds__(1024) void kb_sum_partlog(const BInst *__restrict__ tab,
                                                       const int *__restrict__ list) {
    const BInst &I = tab[list[blockIdx.x]];
    if (!I.do_e) return;
    k_sum_partlog_body(0, 0, 0, 1, I.partLog, I.Np / PT, I.scal);
}

__global__ __launch_bounds__(RK * RD) void kb_reduce32(const BInst *__restrict__ tab,
                                                       const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_e) return;
    const int nkx = (I.K + RK - 1) / RK;
    k_reduce32_body(m.y % nkx, m.y / nkx, 0, 0, I.partRT32, I.partG32, I.NC, I.K, I.D, I.RT,
                    I.G, nullptr);
}

__global__ __launch_bounds__(256) void kb_zero_fix(const BInst *__restrict__ tab,
                                                   const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_e) return;
    k_zero_fix_body(m.y, 0, 0, 0, I.scal, I.T64, I.ldt, I.pm, I.N, I.K, I.D, I.RT, I.G);
}

__global__ __launch_bounds__(NTH) void kb_block_ranges(const BInst *__restrict__ tab,
                                                       const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_e) return;
    block_range_body(m.y, I.bnd, I.Np, I.CL, I.Kp / CT, I.skipBelow, I.brange);
}

__global__ __launch_bounds__(256) void kb_absmax_part(const BInst *__restrict__ tab,
                                                      const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    k_absmax_part_body(m.y, 0, 0, bk_count(I, BK_ABSMAX_PART), I.Y64, (long)I.K, I.ldt, I.D,
                       I.amaxPart);
}

__global__ __launch_bounds__(256) void kb_absmax_final(const BInst *__restrict__ tab,
                                                       const int *__restrict__ list) {
    const BInst &I = tab[list[blockIdx.x]];
    if (!I.do_m) return;
    k_absmax_final_body(0, 0, 0, 1, I.amaxPart, bk_count(I, BK_ABSMAX_PART), I.scal + 6);
}

__global__ __launch_bounds__(256) void kb_make_y16(const BInst *__restrict__ tab,
                                                   const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const BInst &I = tab[m.x];
    if (!I.do_m) return;
    k_make_y16_body(m.y, 0, 0, 0, I.Y64, I.ldt, I.K, I.Kp, I.D, I.ldh, I.scal + 6, I.Y16,
                    I.yy32, I.mscal);
}

// out[i][0 .. kScalCount) = the scalars of table entry i
__global__ void kb_gather_scal(const BInst *__restrict__ tab, int n, double *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * kScalCount) out[e] = tab[e / kScalCount].scal[e % kScalCount];
}

// persistent grid size of a kernel: the workgroups that fit on the device
// (the occupancy query runs once per kernel, before any graph capture uses it)
template <typename F>
int resident_wgs(F kernel, int cus) {
    static const int per = [&] {
        int p = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, kernel, NTH, 0) != hipSuccess ||
            p < 1)
            p = 1;
        return p;
    }();
    return per * cus;
}

template <int KS, int TT>
void launch_e1_batch(hipStream_t st, const BInst *tab, const int *list, int n, int total,
                     int *counter, bool two, int cus) {
    if (two) {
        const int g = std::min(total, resident_wgs(kb_e1_mfma<KS, TT, true>, cus));
        hipLaunchKernelGGL((kb_e1_mfma<KS, TT, true>), dim3((unsigned)g), dim3(NTH), 0, st, tab,
                           list, n, total, counter);
    } else {
        const int g = std::min(total, resident_wgs(kb_e1_mfma<KS, TT, false>, cus));
        hipLaunchKernelGGL((kb_e1_mfma<KS, TT, false>), dim3((unsigned)g), dim3(NTH), 0, st, tab,
                           list, n, total, counter);
    }
}

template <int KS, int TT, bool GM>
void launch_e2_batch_gm(hipStream_t st, const BInst *tab, const int *list, int n, int total,
                        int *counter, bool rlo, int cus) {
    if (rlo) {
        const int g = std::min(total, resident_wgs(kb_e2_mfma<KS, TT, GM, true>, cus));
        hipLaunchKernelGGL((kb_e2_mfma<KS, TT, GM, true>), dim3((unsigned)g), dim3(NTH), 0, st,
                           tab, list, n, total, counter);
    } else {
        const int g = std::min(total, resident_wgs(kb_e2_mfma<KS, TT, GM, false>, cus));
        hipLaunchKernelGGL((kb_e2_mfma<KS, TT, GM, false>), dim3((unsigned)g), dim3(NTH), 0, st,
                           tab, list, n, total, counter);
    }
}

template <int KS, int TT>
void launch_e2_batch(hipStream_t st, const BInst *tab, const int *list, int n, int total,
                     int *counter, bool rlo, int D, int cus) {
    constexpr int DR = 32 * ((16 * KS + 31) / 32);
    if (DR > D) launch_e2_batch_gm<KS, TT, true>(st, tab, list, n, total, counter, rlo, cus);
    else launch_e2_batch_gm<KS, TT, false>(st, tab, list, n, total, counter, rlo, cus);
}

template <int TT>
int launch_class(const BClass &k, hipStream_t st, const BInst *tab, const int *list,
                 int *counters, bool pass2, int cus) {
    const int *l = list + k.first;
    int *ctr = counters + 2 * k.index + (pass2 ? 1 : 0);
#define DG_BATCH_KS(V)                                                                       \
    case V:                                                                                  \
        if (pass2) launch_e2_batch<V, TT>(st, tab, l, k.n, k.items_e2, ctr, k.rlo, k.D, cus); \
        else launch_e1_batch<V, TT>(st, tab, l, k.n, k.items_e1, ctr, k.two, cus);           \
        break;
    switch (k.KS) {
        DG_BATCH_KS(1)
        DG_BATCH_KS(2)
        DG_BATCH_KS(3)
        DG_BATCH_KS(4)
        DG_BATCH_KS(6)
        DG_BATCH_KS(8)
        default:
            if constexpr (TT == 1) {
                switch (k.KS) {
                    DG_BATCH_KS(11)
                    DG_BATCH_KS(12)
                    default: return DG_EINVAL;
                }
            } else {
                return DG_EINVAL;
            }
    }
#undef DG_BATCH_KS
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
        if (b->aK32) (void)hipFree(b->aK32);
        if (b->amaxPart) (void)hipFree(b->amaxPart);
        if (b->gmx) (void)hipFree(b->gmx);
        if (b->bnd) (void)hipFree(b->bnd);
        if (b->brange) (void)hipFree(b->brange);
        if (b->itab) (void)hipFree(b->itab);
        if (b->icnt) (void)hipFree(b->icnt);
        delete b;
    }
    c->mix = nullptr;
    if (c->T16) (void)hipFree(c->T16);
    if (c->tt32) (void)hipFree(c->tt32);
    if (c->yy32) (void)hipFree(c->yy32);
    if (c->pP32) (void)hipFree(c->pP32);
    if (c->partRT32) (void)hipFree(c->partRT32);
    if (c->partG32) (void)hipFree(c->partG32);
    c->T16 = nullptr; c->tt32 = nullptr; c->yy32 = nullptr; c->pP32 = nullptr;
    c->partRT32 = nullptr; c->partG32 = nullptr;
    c->mixed_ready = 0;
}

long dg_mixed_nc_limits(const dg_ctx *c, long *lo, long *hi) {
    const long Np = ((c->N + PT - 1) / PT) * PT;
    const long Kp = ((c->K + CT - 1) / CT) * CT;
    const long blocks = Np / BP;
    // chunks of pass 2: ~8K points each (the FP32 accumulation of R T over a
    // chunk carries a small systematic bias that grows with its length; at 16K
    // points the annealing dynamics amplified it to 3e-6 in llh within four
    // iterations), enough workgroups to fill the chip
    long nmin = (c->N + 8191) / 8192;
    long nc = nmin;
    const long want = (2048 + (Kp / CT) - 1) / (Kp / CT);
    if (nc < want) nc = want;
    if (nc > blocks) nc = blocks;
    if (nc < 1) nc = 1;
    const long top = ((blocks + 7) / 8) * 8;
    if (lo) *lo = std::min(((nmin + 7) / 8) * 8, top);
    if (hi) *hi = top;
    return ((nc + 7) / 8) * 8;          // the pass-2 grid maps 8 chunks per XCD round
}

int dg_mixed_nc(const dg_ctx *c) {
    const MixedBufs *b = reinterpret_cast<const MixedBufs *>(c->mix);
    return b ? b->NC : 0;
}

int dg_mixed_pass2_blocks(dg_ctx *c, long *visited, long *total) {
    MixedBufs *b = mb(c);
    if (!b || !c->mixed_ready) return dg_fail(c, DG_ESTATE, "mixed buffers not prepared");
    const int KG = b->Kp / CT;
    std::vector<int2> rg((size_t)KG * b->NC * RS);
    DG_HIP(c, hipMemcpyAsync(rg.data(), b->brange, sizeof(int2) * rg.size(), hipMemcpyDeviceToHost,
                             c->stream));
    DG_HIP(c, hipStreamSynchronize(c->stream));
    long v = 0;
    for (size_t i = 0; i < rg.size(); i += RS) v += rg[i].y;
    *visited = v;
    *total = (long)KG * (b->Np / BP);
    return DG_OK;
}

// T16 (both FP16 terms, rows in the mixed path's point order), tt32 and the
// chunk-major transposed copy T16t
int build_t16(dg_ctx *c, MixedBufs *b) {
    const int ld16 = c->ldh;
    const int DR = 32 * ((ld16 + 31) / 32);
    hipLaunchKernelGGL(k_make_t16, dim3((unsigned)((b->Np + 255) / 256)), dim3(256), 0, c->stream,
                       c->T64, c->ldt, c->N, b->Np, c->D, ld16, b->TT, (double)c->t_scale,
                       c->perm, (_Float16 *)c->T16, c->tt32);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_transpose16,
                       dim3((unsigned)((b->Np + 63) / 64), (unsigned)((DR + 63) / 64),
                            (unsigned)b->TT),
                       dim3(256), 0, c->stream, (const _Float16 *)c->T16, c->N, b->Np, ld16, DR,
                       DR > c->D ? c->D : -1, b->CL, b->TT, b->T16t);
    DG_HIP(c, hipGetLastError());
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
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
    // pass-2 chunk count: the context's override (a tuned value) within the
    // accuracy and grid limits, else the built-in choice
    long lo = 0, hi = 0;
    long nc = dg_mixed_nc_limits(c, &lo, &hi);
    if (c->nc_override > 0) nc = std::max(lo, std::min(hi, ((long)c->nc_override + 7) / 8 * 8));
    b->NC = (int)nc;
    // chunk length: a multiple of 64 points (transpose tiles never straddle chunks)
    b->CL = ((b->Np + nc * 64 - 1) / (nc * 64)) * 64;

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
    // pass 2 addresses the data rows with 32-bit byte offsets
    if ((double)b->TT * (double)b->Np * ld16 * 2.0 >= 4294967296.0)
        return dg_fail(c, DG_EINVAL, "mixed path: N x D too large for one context (split N)");
    double tscale = 1.0;
    if (!c->t_exact && amax > 0.0) tscale = ldexp(1.0, 13 - ilogb(amax));
    c->t_scale = (float)tscale;

    if ((rc = dalloc(c, (_Float16 **)&c->T16, (size_t)b->TT * b->Np * ld16)) ||
        (rc = dalloc(c, &b->T16t, (size_t)b->NC * b->TT * DR * b->CL)) ||
        (rc = dalloc(c, &c->tt32, (size_t)b->Np)) ||
        (rc = dalloc(c, &b->Y16, (size_t)2 * b->Kp * ld16)) ||
        (rc = dalloc(c, &c->yy32, (size_t)b->Kp)) || (rc = dalloc(c, &b->mscal, 4)) ||
        (rc = dalloc(c, &b->partLog, (size_t)3 * (b->Np / PT))) ||
        (rc = dalloc(c, &b->blkThr, (size_t)(b->Np / BP))) ||
        (rc = dalloc(c, &b->aK32, (size_t)(b->Kp / KT) * AKS)) ||
        (rc = dalloc(c, &c->pP32, (size_t)2 * b->Np)) ||
        (rc = dalloc(c, &c->partRT32, (size_t)b->NC * c->K * c->D)) ||
        (rc = dalloc(c, &c->partG32, (size_t)b->NC * c->K)))
        return rc;
    // pass-2 block skipping: the group maxima and bounds pass 1 leaves
    const size_t KG = (size_t)(b->Kp / CT);
    if ((rc = dalloc(c, &b->brange, (size_t)KG * b->NC * RS)) ||
        (rc = dalloc(c, &b->itab, 8 * KG * ((b->NC + 7) / 8))) || (rc = dalloc(c, &b->icnt, 8)))
        return rc;
    b->screen = !env_off("DG_MIXED_SCREEN");
    if (!env_off("DG_MIXED_SKIP") &&
        ((rc = dalloc(c, &b->gmx, KG * b->Np)) || (rc = dalloc(c, &b->bnd, KG * (b->Np / BP)))))
        return rc;
    double ms[4] = {1.0, tscale, 0.0, 0.0};
    DG_HIP(c, hipMemcpyAsync(b->mscal, ms, sizeof(ms), hipMemcpyHostToDevice, c->stream));
    int rcb = build_t16(c, b);
    if (rcb) return rcb;
    c->mixed_ready = 1;
    return DG_OK;
}

// Point ordering: the points are ordered by their nearest centre under the
// current centres (k_nearest; a stable counting sort, input order among
// points with the same centre) and the FP16 operands rebuilt in that order.
// Points with the same nearest centre have their responsibilities on the same
// centres, so pass 2 finds whole (centre group, block) pairs with nothing to
// add.  Any order gives the same results up to the order of the sums.
int dg_mixed_order(dg_ctx *c) {
    MixedBufs *b = mb(c);
    if (!b || !c->mixed_ready) return dg_fail(c, DG_ESTATE, "mixed buffers not prepared");
    if (c->perm || env_off("DG_MIXED_ORDER")) return DG_OK;
    const long N = c->N, Np = b->Np;
    const int K = c->K;
    int *dkey = nullptr;
    int rc;
    if ((rc = dalloc(c, &dkey, (size_t)N))) return rc;
    const unsigned grid = (unsigned)((N + 255) / 256);
    auto go = [&](auto ks) {
        constexpr int KS = decltype(ks)::value;
        hipLaunchKernelGGL((k_nearest<KS>), dim3(grid), dim3(256), 0, c->stream,
                           (const _Float16 *)c->T16, N, b->Y16, c->yy32, K, b->mscal, dkey);
    };
    switch (b->KS) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 6: go(std::integral_constant<int, 6>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 11: go(std::integral_constant<int, 11>{}); break;
        default: go(std::integral_constant<int, 12>{}); break;
    }
    std::vector<int> key((size_t)N), perm((size_t)Np);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(key.data(), dkey, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dkey);
    if (e != hipSuccess) return dg_hip_check(c, e, "point ordering");
    std::vector<long> start((size_t)K + 1, 0);
    for (long n = 0; n < N; ++n) {
        const int k = key[(size_t)n];
        if (k < 0 || k >= K) return dg_fail(c, DG_EHIP, "point ordering: bad centre index");
        ++start[(size_t)k + 1];
    }
    for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
    for (long n = 0; n < N; ++n) perm[(size_t)start[(size_t)key[(size_t)n]]++] = (int)n;
    for (long n = N; n < Np; ++n) perm[(size_t)n] = (int)n;
    if ((rc = dalloc(c, &c->perm, (size_t)Np))) return rc;
    DG_HIP(c, hipMemcpyAsync(c->perm, perm.data(), sizeof(int) * Np, hipMemcpyHostToDevice,
                             c->stream));
    // each pass-1 workgroup starts its sweep at the centre group of its middle
    // point's nearest centre, so that its running maxima are near their final
    // values early and the screen leaves out most of the other tiles
    const long nwg = Np / PT;
    std::vector<int> st((size_t)nwg);
    for (long w = 0; w < nwg; ++w) {
        const long n = std::min(N - 1, w * PT + PT / 2);
        st[(size_t)w] = key[(size_t)perm[(size_t)n]] / CT * (CT / KT);
    }
    if ((rc = dalloc(c, &c->p1start, (size_t)nwg))) return rc;
    DG_HIP(c, hipMemcpyAsync(c->p1start, st.data(), sizeof(int) * nwg, hipMemcpyHostToDevice,
                             c->stream));
    DG_HIP(c, hipStreamSynchronize(c->stream));    // the host vectors go out of scope
    return build_t16(c, b);
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
    hipLaunchKernelGGL(k_make_y16, dim3((unsigned)((b->Kp + 3) / 4)), dim3(256), 0, c->stream,
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
    hipLaunchKernelGGL(k_reduce32,
                       dim3((unsigned)((c->K + RK - 1) / RK), (unsigned)((c->D + RD - 1) / RD)),
                       dim3(RK * RD), 0, c->stream, c->partRT32, c->partG32, b->NC, c->K, c->D,
                       c->RT, c->G, b->brange);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_sum_partlog, dim3(1), dim3(1024), 0, c->stream, b->partLog, b->Np / PT,
                       c->scal);
    DG_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_zero_fix, dim3((unsigned)c->D), dim3(256), 0, c->stream, c->scal, c->T64,
                       c->ldt, c->pm, c->N, c->K, c->D, c->RT, c->G);
    DG_HIP(c, hipGetLastError());
    return DG_OK;
}

// ---- batched E-step ---------------------------------------------------------
int dg_batch_fill_mixed(const dg_ctx *c, BInst *I) {
    const MixedBufs *b = reinterpret_cast<const MixedBufs *>(c->mix);
    if (!b || !c->mixed_ready) return DG_ESTATE;
    I->Np = b->Np; I->CL = b->CL; I->Kp = b->Kp; I->ldh = c->ldh; I->NC = b->NC; I->KS = b->KS;
    I->TT = b->TT;
    I->T16 = (const _Float16 *)c->T16; I->T16t = b->T16t; I->Y16 = b->Y16;
    I->tt32 = c->tt32; I->yy32 = c->yy32; I->aK32 = b->aK32; I->pP32 = c->pP32;
    I->blkThr = b->blkThr; I->partRT32 = c->partRT32; I->partG32 = c->partG32;
    // the batched pass 1 keeps no group maxima for KS = 8 with split data
    const bool gmx_ok = !(b->KS == 8 && b->TT == 2);
    I->gmx = gmx_ok ? b->gmx : nullptr;
    I->bnd = gmx_ok ? b->bnd : nullptr;
    I->perm = c->perm;
    I->p1start = c->p1start;
    I->screen = b->screen ? 1 : 0;
    I->brange = b->brange;
    I->skipBelow = skip_below(32 * ((16 * b->KS + 31) / 32) > c->D);
    I->mscal = b->mscal; I->partLog = b->partLog; I->amaxPart = b->amaxPart;
    return DG_OK;
}

int dg_batch_items(const BInst &I, int *e1, int *e2) {
    *e1 = (int)(I.Np / PT);
    *e2 = (I.Kp / CT) * I.NC;
    return DG_OK;
}

int dg_batch_estep(hipStream_t st, const BInst *tab, const int *list, int n, const BClass *cls,
                   int ncls, int *counters, const BGrids &gr, int cus, hipStream_t *side,
                   int nside, hipEvent_t *ev) {
    if (n <= 0) return DG_OK;
    hipLaunchKernelGGL(kb_zero_ints, dim3((unsigned)((4 * ncls + 255) / 256)), dim3(256), 0, st,
                       counters, 4 * ncls);
    hipLaunchKernelGGL(kb_make_ak, dim3((unsigned)gr.wg[BK_MAKE_AK]), dim3(256), 0, st, tab,
                       gr.map[BK_MAKE_AK]);
    // the classes are independent until the scalars are gathered: class i runs
    // on side stream i % nside (fork after make_ak, join before the gather)
    if (hipEventRecord(ev[0], st) != hipSuccess) return DG_EHIP;
    for (int j = 0; j < nside; ++j)
        if (hipStreamWaitEvent(side[j], ev[0], 0) != hipSuccess) return DG_EHIP;
    int rc;
    for (int i = 0; i < ncls; ++i) {
        const BClass &k = cls[i];
        hipStream_t ss = side[i % nside];
        const int *l = list + k.first;
        const unsigned un = (unsigned)k.n;
        rc = (k.TT == 1) ? launch_class<1>(k, ss, tab, list, counters, false, cus)
                         : launch_class<2>(k, ss, tab, list, counters, false, cus);
        if (rc) return rc;
        hipLaunchKernelGGL(kb_sum_partlog, dim3(un), dim3(1024), 0, ss, tab, l);
        hipLaunchKernelGGL(kb_block_ranges, dim3((unsigned)k.wg_lists), dim3(NTH), 0, ss, tab,
                           k.map_lists);
        rc = (k.TT == 1) ? launch_class<1>(k, ss, tab, list, counters, true, cus)
                         : launch_class<2>(k, ss, tab, list, counters, true, cus);
        if (rc) return rc;
        hipLaunchKernelGGL(kb_reduce32, dim3((unsigned)k.wg_reduce32), dim3(RK * RD), 0, ss, tab,
                           k.map_reduce32);
        hipLaunchKernelGGL(kb_zero_fix, dim3((unsigned)k.wg_zero_fix), dim3(256), 0, ss, tab,
                           k.map_zero_fix);
    }
    for (int j = 0; j < nside; ++j) {
        if (hipEventRecord(ev[1 + j], side[j]) != hipSuccess) return DG_EHIP;
        if (hipStreamWaitEvent(st, ev[1 + j], 0) != hipSuccess) return DG_EHIP;
    }
    return hipGetLastError() == hipSuccess ? DG_OK : DG_EHIP;
}

int dg_batch_prepare_y(hipStream_t st, const BInst *tab, const int *list, int n,
                       const BGrids &gr) {
    if (n <= 0) return DG_OK;
    hipLaunchKernelGGL(kb_absmax_part, dim3((unsigned)gr.wg[BK_ABSMAX_PART]), dim3(256), 0, st, tab,
                       gr.map[BK_ABSMAX_PART]);
    hipLaunchKernelGGL(kb_absmax_final, dim3((unsigned)n), dim3(256), 0, st, tab, list);
    hipLaunchKernelGGL(kb_make_y16, dim3((unsigned)gr.wg[BK_MAKE_Y16]), dim3(256), 0, st, tab,
                       gr.map[BK_MAKE_Y16]);
    return hipGetLastError() == hipSuccess ? DG_OK : DG_EHIP;
}

int dg_batch_gather_scal(hipStream_t st, const BInst *tab, int n, double *out) {
    if (n <= 0) return DG_OK;
    hipLaunchKernelGGL(kb_gather_scal, dim3((unsigned)((n * kScalCount + 255) / 256)), dim3(256), 0,
                       st, tab, n, out);
    return hipGetLastError() == hipSuccess ? DG_OK : DG_EHIP;
}
