// Mixed-precision Poisson solve (precision = BL_PREC_MIXED).
//
// The pseudo-pressure rhs (projection.f90:34-49) is formed in FP32; the x
// cosine transform and the z Fourier transform run as batched GEMMs on the
// matrix cores, each product split into BF16 pieces (a = a_hi + a_lo):
//   a*b ~ a_hi*b_hi + a_hi*b_lo + a_lo*b_hi, accumulated in FP32;
// the tridiagonal solves in y use the FP64 LU factors of
// initialization.f90:427-469 with FP64 arithmetic.  The near-singular
// (kx, kz) = (0, 0) mode bypasses the GEMMs: its rhs is an FP64 sum of the
// divergence over each x-z plane, it is solved in FP64, and its constant
// contribution phi00(j) / (2 N M) is added back after the inverse x GEMM.
//
// Transform matrices (N = nx-2, M = nz-2), identical in meaning to the FP64
// path's DCT/DFT (projection.f90:110-202 with the 1/(2 nxp nzp) factor):
//   Dxf[k][n] = 2 cos(pi k (2n+1) / (2N))
//   Dxi[n][k] = c_k cos(pi k (2n+1) / (2N)) / (N M),  c_0 = 1/2, c_k = 1
//   Zf[q][m]: q = 0 -> 1, q = 2k-1 -> cos(2 pi k m / M), q = 2k -> -sin(2 pi k m / M)
//   Zi[m][q]: q = 0 -> 1, q = 2k-1 -> 2 cos (cos at k = M/2), q = 2k -> -2 sin
// Spectral work arrays: T1[m][j][kx] and S[q][j][kx], kx stride NP = N
// rounded up to 32 (the padding stays zero).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "bl_internal.hpp"

namespace blgpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short bf16_bits(float x) {
    const __bf16 b = (__bf16)x;  // round to nearest even
    return __builtin_bit_cast(unsigned short, b);
}
__device__ __forceinline__ float bf16_float(unsigned short u) {
    return __builtin_bit_cast(float, (unsigned int)u << 16);
}

// ------------------------------------------------------------------------
// C[r][c] = sum_k A[r][k] B[k][c] (+ add[c % o_C1]) with A pre-split into
// BF16 hi/lo rows of length Kp (zero padded past K, rows padded to a whole
// number of BM tiles) and B split on the fly (b = b_hi + b_lo, FP32 in HBM).
// Workgroup tile BM x BN = (64 WR) x (64 WC), one 64 x 64 tile per wave,
// K step 32.  MFMA v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+e],
// B[8(l/16)+e][l%16] (e < 8) and C[4(l/16)+r][l%16] (r < 4).
// Software pipeline: the B tile of step s+1 is loaded into registers and the
// A fragments of step s+1 are fetched while step s's MFMAs run out of one
// LDS buffer; the converted B tile then goes to the other buffer (one
// barrier per step).
// B layout: KCONTIG -> B[c*ldb + k] (needs Kp <= ldb: the row tail past K is
// read and masked), else B[k*ldb + c] (needs C % 4 == 0, ldb % 4 == 0).
// Output layout: RCONTIG -> out[o_base + (c%o_C1)*o_c1 + (c/o_C1)*o_c2 + r],
// else out[r*ldc + c].
// ------------------------------------------------------------------------
struct GemmArgs {
    int R, C, K, Kp;
    const unsigned short *Ahi, *Alo;
    const float *B;
    long long ldb;
    float *out;
    long long o_base, o_c1, o_c2, ldc;
    int o_C1;
    const double *add;  // optional per-(c % o_C1) addend
};

// K step BK (a multiple of 32); an LDS row of one column holds BK + 8 values
constexpr int kGemmLds = 163840;  // LDS bytes per workgroup on gfx950

// Offset of column col of a B tile in LDS (bf16 units): rows of LDK = BK + 8
// plus 16 bytes per group of 16 columns.  The transposed stores of a wave
// (lane cg writes columns 4 cg .. 4 cg + 3) then meet each bank at most twice
// instead of eight times (64 banks of 4 bytes), while the 16 columns that one
// fragment read covers share one skew and stay conflict-free; each column's
// k run stays 16-byte aligned and the runs do not overlap.
template <int BK>
__device__ __forceinline__ constexpr int gemm_bcol(int col) {
    return col * (BK + 8) + (col >> 4) * 8;
}
template <int WC, int BK>
constexpr int gemm_bsize() {
    return 64 * WC * (BK + 8) + 4 * WC * 8;
}

// float4 loads of the B tile per thread
template <int WR, int WC, int BK>
constexpr int gemm_bloads() {
    return (BK * 64 * WC / 4) / (64 * WR * WC);
}

// B tile loads: unconditional (addresses clamped into the operand) so that
// the loads of the next step stay in flight across the MFMAs; the masks
// (columns past C, k past K) are applied when the tile is split and stored.
template <int WR, int WC, int BK, bool KCONTIG>
__device__ __forceinline__ void gemm_load_b(const GemmArgs &g, int c0, int kk, int tid,
                                            float4 (&v)[gemm_bloads<WR, WC, BK>()]) {
    constexpr int NT = 64 * WR * WC, BN = 64 * WC, NL = gemm_bloads<WR, WC, BK>();
    if constexpr (KCONTIG) {
        // BK / 4 threads cover the BK k of one column
        constexpr int TPC = BK / 4;
#pragma unroll
        for (int p = 0; p < NL; ++p) {
            const int c = min(c0 + p * (NT / TPC) + tid / TPC, g.C - 1), k = kk + 4 * (tid % TPC);
            v[p] = *reinterpret_cast<const float4 *>(g.B + (long long)c * g.ldb + k);
        }
    } else {
        // 4 consecutive columns per thread, BN / 4 threads per k row
        static_assert(NL % 2 == 0, "k pairs");
        const int kq = tid / (BN / 4), cg = tid % (BN / 4);
        const int c = min(c0 + 4 * cg, g.C - 4);
#pragma unroll
        for (int p = 0; p < NL; ++p) {
            const int k = min(kk + kq * NL + p, g.K - 1);
            v[p] = *reinterpret_cast<const float4 *>(g.B + (long long)k * g.ldb + c);
        }
    }
}

__device__ __forceinline__ void split2(float x, unsigned short &h, unsigned short &l) {
    h = bf16_bits(x);
    l = bf16_bits(x - bf16_float(h));
}

template <int WR, int WC, int BK, bool KCONTIG>
__device__ __forceinline__ void gemm_store_b(const GemmArgs &g, int c0, int kk, int tid,
                                             const float4 (&v)[gemm_bloads<WR, WC, BK>()], unsigned short *sBh,
                                             unsigned short *sBl) {
    constexpr int NT = 64 * WR * WC, BN = 64 * WC, NL = gemm_bloads<WR, WC, BK>();
    if constexpr (KCONTIG) {
        constexpr int TPC = BK / 4;
        const int k = 4 * (tid % TPC);
        const int kg = kk + k;
#pragma unroll
        for (int p = 0; p < NL; ++p) {
            const int col = p * (NT / TPC) + tid / TPC;
            const bool cin = c0 + col < g.C;
            const float e[4] = {cin && kg < g.K ? v[p].x : 0.f, cin && kg + 1 < g.K ? v[p].y : 0.f,
                                cin && kg + 2 < g.K ? v[p].z : 0.f, cin && kg + 3 < g.K ? v[p].w : 0.f};
            unsigned short h[4], l[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) split2(e[q], h[q], l[q]);
            const int o = gemm_bcol<BK>(col) + k;
            *reinterpret_cast<uint2 *>(&sBh[o]) =
                make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16));
            *reinterpret_cast<uint2 *>(&sBl[o]) =
                make_uint2((unsigned)l[0] | ((unsigned)l[1] << 16), (unsigned)l[2] | ((unsigned)l[3] << 16));
        }
    } else {
        const int kq = tid / (BN / 4), cg = tid % (BN / 4);
        const bool cin = c0 + 4 * cg < g.C;
#pragma unroll
        for (int p = 0; p < NL; p += 2) {
            const int k = kq * NL + p;
            const bool k0 = cin && kk + k < g.K, k1 = cin && kk + k + 1 < g.K;
            const float a0[4] = {v[p].x, v[p].y, v[p].z, v[p].w};
            const float a1[4] = {v[p + 1].x, v[p + 1].y, v[p + 1].z, v[p + 1].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                unsigned short h0, l0, h1, l1;
                split2(k0 ? a0[e] : 0.f, h0, l0);
                split2(k1 ? a1[e] : 0.f, h1, l1);
                const int o = gemm_bcol<BK>(4 * cg + e) + k;
                *reinterpret_cast<unsigned int *>(&sBh[o]) = (unsigned)h0 | ((unsigned)h1 << 16);
                *reinterpret_cast<unsigned int *>(&sBl[o]) = (unsigned)l0 | ((unsigned)l1 << 16);
            }
        }
    }
}

// A tile loads: BM rows x BK k of hi and of lo as 16-byte chunks of 8 k
// (rows are padded to whole tiles and k to Kp, so no masks)
template <int WR, int WC, int BK>
constexpr int gemm_aloads() {
    return BK / (4 * WC);  // (BM * BK / 8 chunks * 2) / NT
}

template <int WR, int WC, int BK>
__device__ __forceinline__ void gemm_load_a(const GemmArgs &g, int r0, int kk, int tid,
                                            u16x8 (&va)[gemm_aloads<WR, WC, BK>()]) {
    constexpr int NT = 64 * WR * WC, NA = gemm_aloads<WR, WC, BK>(), kCPR = BK / 8;
#pragma unroll
    for (int p = 0; p < NA; ++p) {
        const int q = p * NT + tid;  // chunk id over [hi, lo][row][kCPR chunks]
        const int half = q / (64 * WR * kCPR), rc = q % (64 * WR * kCPR);
        const int row = rc / kCPR, ch = rc % kCPR;
        const unsigned short *src = (half ? g.Alo : g.Ahi) + (long long)(r0 + row) * g.Kp + kk + 8 * ch;
        va[p] = *reinterpret_cast<const u16x8 *>(src);
    }
}

template <int WR, int WC, int BK>
__device__ __forceinline__ void gemm_store_a(int tid, const u16x8 (&va)[gemm_aloads<WR, WC, BK>()],
                                             unsigned short *sAh, unsigned short *sAl) {
    constexpr int NT = 64 * WR * WC, NA = gemm_aloads<WR, WC, BK>(), kCPR = BK / 8, LDK = BK + 8;
#pragma unroll
    for (int p = 0; p < NA; ++p) {
        const int q = p * NT + tid;
        const int half = q / (64 * WR * kCPR), rc = q % (64 * WR * kCPR);
        const int row = rc / kCPR, ch = rc % kCPR;
        *reinterpret_cast<u16x8 *>((half ? sAl : sAh) + row * LDK + 8 * ch) = va[p];
    }
}

template <int WR, int WC, int BK, bool KCONTIG, bool RCONTIG>
__global__ __launch_bounds__(64 * WR * WC) void k_gemm_bf16x3(GemmArgs g, int tiles_r, int nblk) {
    constexpr int BM = 64 * WR, BN = 64 * WC, NL = gemm_bloads<WR, WC, BK>(), NA = gemm_aloads<WR, WC, BK>();
    constexpr int kLDK = BK + 8;
    static_assert(BK % 32 == 0, "K step");
    // one buffer of each operand: [row or column][k], rows padded to kLDK
    __shared__ unsigned short sA[2][BM * kLDK], sB[2][gemm_bsize<WC, BK>()];  // [hi, lo]
    static_assert(sizeof(unsigned short) * 2 * (BM * kLDK + gemm_bsize<WC, BK>()) <= kGemmLds, "LDS per workgroup");
    // XCD-aware order: blocks that share a column tile run on one XCD (the
    // hardware deals consecutive block ids round-robin over 8 XCDs)
    const int b = blockIdx.x;
    const int per = nblk / 8;
    const int lb = (b % 8) < (nblk % 8) ? (b % 8) * (per + 1) + b / 8
                                        : (nblk % 8) * (per + 1) + ((b % 8) - (nblk % 8)) * per + b / 8;
    const int tr = lb % tiles_r, tc = lb / tiles_r;
    const int r0 = tr * BM, c0 = tc * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave % WR) * 64, wc = (wave / WR) * 64;
    const int l16 = lane & 15, lq = lane >> 4;

    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    float4 vb[NL];
    u16x8 va[NA];
    gemm_load_b<WR, WC, BK, KCONTIG>(g, c0, 0, tid, vb);
    gemm_load_a<WR, WC, BK>(g, r0, 0, tid, va);
    gemm_store_b<WR, WC, BK, KCONTIG>(g, c0, 0, tid, vb, sB[0], sB[1]);
    gemm_store_a<WR, WC, BK>(tid, va, sA[0], sA[1]);
    __syncthreads();
    for (int kk = 0; kk < g.Kp; kk += BK) {
        const bool more = kk + BK < g.Kp;
        // the next step's tiles stay in flight during this step's MFMAs
        if (more) {
            gemm_load_b<WR, WC, BK, KCONTIG>(g, c0, kk + BK, tid, vb);
            gemm_load_a<WR, WC, BK>(g, r0, kk + BK, tid, va);
        }
#pragma unroll
        for (int ks = 0; ks < BK / 32; ++ks) {
            bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int oa = (wr + q * 16 + l16) * kLDK + 32 * ks + 8 * lq;
                const int ob = gemm_bcol<BK>(wc + q * 16 + l16) + 32 * ks + 8 * lq;
                ah[q] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&sA[0][oa]));
                al[q] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&sA[1][oa]));
                bh[q] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&sB[0][ob]));
                bl[q] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&sB[1][ob]));
            }
            // ---- 4 x 4 blocks of 16 x 16, three products each (small terms
            // first), 16 independent accumulators between dependent MFMAs
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], bh[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bl[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bh[cb], acc[rb][cb], 0, 0, 0);
        }
        if (more) {
            __syncthreads();  // every wave has read this step's tiles
            gemm_store_b<WR, WC, BK, KCONTIG>(g, c0, kk + BK, tid, vb, sB[0], sB[1]);
            gemm_store_a<WR, WC, BK>(tid, va, sA[0], sA[1]);
            __syncthreads();
        }
    }
    // ---- epilogue
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int c = c0 + wc + cb * 16 + l16;
        if (c >= g.C) continue;
        const float addv = g.add ? (float)g.add[c % g.o_C1] : 0.f;
        const long long ob = RCONTIG ? g.o_base + (long long)(c % g.o_C1) * g.o_c1 + (long long)(c / g.o_C1) * g.o_c2
                                     : (long long)c;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + wr + rb * 16 + 4 * lq + q;
                if (r >= g.R) continue;
                const float v = acc[rb][cb][q] + addv;
                if (RCONTIG)
                    g.out[ob + r] = v;
                else
                    g.out[(long long)r * g.ldc + ob] = v;
            }
    }
}

void check();

template <int WR, int WC, int BK, bool KCONTIG, bool RCONTIG>
void gemm(Handle &h, int kid, const GemmArgs &g) {
    ProfScope ps(h, kid);
    constexpr int BM = 64 * WR, BN = 64 * WC;
    // the loaders' assumptions (see k_gemm_bf16x3)
    const bool ok = g.Kp % BK == 0 && g.Kp >= g.K &&
                    g.C >= 4 && g.K >= 1 &&
                    (KCONTIG ? (g.Kp <= g.ldb && g.ldb % 4 == 0) : (g.C % 4 == 0 && g.ldb % 4 == 0 && g.ldb >= g.C));
    if (!ok) {
        last_error() = "internal: GEMM operand layout does not fit the kernel";
        throw DeviceError();
    }
    const int tr = (g.R + BM - 1) / BM, tc = (g.C + BN - 1) / BN;
    const int nblk = tr * tc;
    hipLaunchKernelGGL((k_gemm_bf16x3<WR, WC, BK, KCONTIG, RCONTIG>), dim3(nblk), dim3(64 * WR * WC), 0, h.stream, g, tr,
                       nblk);
    check();
}

// ------------------------------------------------------------------------
// The real z DFT (M points, M/2 + 1 <= 96) folded by the symmetry m <-> M-m,
// as the FP64 kernel k_zdft_mf does, with split-BF16x3 products accumulated
// in FP32: D = A B over the folded index kk = 0..M/2, A[row][kk] = cos or
// sin(2 pi row kk / M) pre-split on the host, B the folded pairs split on the
// fly (forward: e, o = T[kk] +- T[M-kk]; inverse: the Hermitian-weighted Re
// and Im rows R, I).  Half the products of the dense 191 x 191 GEMM.
// v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+e], B[8(l/16)+e][l%16]
// (e < 8) and D[4(l/16)+r][l%16] (r < 4).  A block holds every A fragment in
// LDS (cos hi, cos lo, sin hi, sin lo per lane, RT x 3 k steps: 72 KB at
// M = 191) and its waves loop over 16-column tiles; the raw inputs of the
// next k step are loaded as each one's are consumed.
// ------------------------------------------------------------------------
constexpr int kZxRT = 6, kZxKS = 3, kZxW = 8;

template <bool FWD, int RT>
__global__ __launch_bounds__(64 * kZxW) void k_zdft_mx(int M, int C, long long plane, const u16x8 *__restrict__ zfr,
                                                       const float *__restrict__ in, float *__restrict__ out) {
    __shared__ u16x8 sA[RT * kZxKS * 64 * 4];
    const int hm = M / 2, Kc = hm + 1, KS = (Kc + 31) / 32;
    for (int e = threadIdx.x; e < RT * KS * 64 * 4; e += 64 * kZxW) sA[e] = zfr[e];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, c16 = lane & 15;
    const int ntile = (C + 15) / 16, stride = gridDim.x * kZxW;
    int tile = blockIdx.x * kZxW + wave;
    if (tile >= ntile) return;
    // raw inputs of k step s of tile tl: kk = 32 s + 8 q + e (rows clamped
    // to M/2, columns to C - 1; masked where used), one step ahead: the
    // steps run (tile, s) in order, the last step's successor is the next
    // tile's first.  Offsets are 32-bit (M plane < 2^31, checked on the host).
    float pa[8], pb[8];
    auto load = [&](int tl, int s) {
        const float *src = in + min(16 * tl + c16, C - 1);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int kk = min(32 * s + 8 * q + e, hm);
            const bool single = kk == 0 || 2 * kk == M;
            int r0, r1;
            if (FWD) {
                r0 = kk;
                r1 = single ? kk : M - kk;
            } else {
                r0 = kk == 0 ? 0 : (2 * kk == M ? M - 1 : 2 * kk - 1);
                r1 = single ? r0 : 2 * kk;
            }
            pa[e] = src[r0 * (int)plane];
            pb[e] = src[r1 * (int)plane];
        }
    };
    load(tile, 0);
    for (; tile < ntile; tile += stride) {
        f32x4 ac[RT], as[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            ac[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            as[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        for (int s = 0; s < KS; ++s) {
            u16x8 bch, bcl, bsh, bsl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = 32 * s + 8 * q + e;
                const bool live = kk <= hm, single = kk == 0 || 2 * kk == M;
                const float a = live ? pa[e] : 0.f, b = live && !single ? pb[e] : 0.f;
                float vc, vs;
                if (FWD) {
                    vc = a + b;
                    vs = single ? 0.f : a - b;
                } else {
                    vc = single ? a : 2.f * a;
                    vs = 2.f * b;
                }
                unsigned short h0, l0, h1, l1;
                split2(vc, h0, l0);
                split2(vs, h1, l1);
                bch[e] = h0;
                bcl[e] = l0;
                bsh[e] = h1;
                bsl[e] = l1;
            }
            const bool here = s + 1 < KS;
            load(here ? tile : min(tile + stride, ntile - 1), here ? s + 1 : 0);
            const bf16x8 ch = __builtin_bit_cast(bf16x8, bch), cl = __builtin_bit_cast(bf16x8, bcl);
            const bf16x8 sh = __builtin_bit_cast(bf16x8, bsh), sl = __builtin_bit_cast(bf16x8, bsl);
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const u16x8 *f = sA + ((t * KS + s) * 64 + lane) * 4;
                const bf16x8 fch = __builtin_bit_cast(bf16x8, f[0]), fcl = __builtin_bit_cast(bf16x8, f[1]);
                const bf16x8 fsh = __builtin_bit_cast(bf16x8, f[2]), fsl = __builtin_bit_cast(bf16x8, f[3]);
                // small terms first, as in k_gemm_bf16x3
                ac[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fcl, ch, ac[t], 0, 0, 0);
                ac[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fch, cl, ac[t], 0, 0, 0);
                ac[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fch, ch, ac[t], 0, 0, 0);
                as[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fsl, sh, as[t], 0, 0, 0);
                as[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fsh, sl, as[t], 0, 0, 0);
                as[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fsh, sh, as[t], 0, 0, 0);
            }
        }
        const int col = 16 * tile + c16;
        if (col < C) {
            float *dst = out + col;
#pragma unroll
            for (int t = 0; t < RT; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * t + 4 * q + r;
                    if (row > hm) continue;
                    const float c = ac[t][r], sv = as[t][r];
                    if (FWD) {
                        dst[(long long)(row == 0 ? 0 : 2 * row - 1) * plane] = c;
                        if (row >= 1 && 2 * row <= M - 1) dst[(long long)(2 * row) * plane] = -sv;
                    } else {
                        dst[(long long)row * plane] = c - sv;
                        if (row >= 1 && M - row != row) dst[(long long)(M - row) * plane] = c + sv;
                    }
                }
            }
        }
    }
}

// the folded z DFT with RT = rt row tiles; persistent blocks, as many as fit
template <int RT>
void zdft_mx_launch(Handle &h, int kid, int rt, int M, int C, long long plane, bool forward, const float *in,
                    float *out) {
    if constexpr (RT > 1) {
        if (rt < RT) {
            zdft_mx_launch<RT - 1>(h, kid, rt, M, C, plane, forward, in, out);
            return;
        }
    }
    if ((long long)M * plane >= (1LL << 31)) {
        last_error() = "z transform: M x plane exceeds 32-bit offsets";
        throw DeviceError();
    }
    ProfScope ps(h, kid);
    const int tiles = (C + 15) / 16;
    int per_cu = 0;
    if (forward)
        BL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_zdft_mx<true, RT>, 64 * kZxW, 0));
    else
        BL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_zdft_mx<false, RT>, 64 * kZxW, 0));
    const dim3 grid(
        (unsigned)std::max(1, std::min(std::max(1, per_cu) * std::max(1, h.n_cu), (tiles + kZxW - 1) / kZxW)));
    const u16x8 *zfr = reinterpret_cast<const u16x8 *>(h.zmf);
    if (forward)
        hipLaunchKernelGGL((k_zdft_mx<true, RT>), grid, dim3(64 * kZxW), 0, h.stream, M, C, plane, zfr, in, out);
    else
        hipLaunchKernelGGL((k_zdft_mx<false, RT>), grid, dim3(64 * kZxW), 0, h.stream, M, C, plane, zfr, in, out);
    check();
}

// ------------------------------------------------------------------------
// x transform as two short DFTs on the matrix cores (N = N1 N2, Makhoul's
// reordering, k1 <= N1/2 only; the same factorisation as the FP64 kernels),
// each DFT a small GEMM with split-BF16x3 products accumulated in FP32.
// Forward, step A: [Re, Im of B[k1][n2]] = WA [n1 x (line, n2)], K = N1 <= 32
//   (one K step); step C: [Re, Im of V[k1 + N1 k2]] = WC [(n2, re/im) x
//   (line, k1)], K = 2 N2 <= 64 (two K steps).  Inverse: the conjugate DFT
//   over k2 (step C'), then the real output over (k1, re/im) with the
//   Hermitian weights (step A').
// Operand tables (host, lane order): per (row tile, K step) and lane l the
// BF16 hi and lo of A[l%16][8(l/16) + e], e < 8.  In a paired table a tile
// row R is (Re, Im) = R % 2 of index 8 rt + 2 (R/4) + (R%4)/2, so that a
// lane's four accumulator rows 4(l/16) + r hold (Re, Im) of two indices.
// ------------------------------------------------------------------------
struct XArgs {
    int N, N1, N2, K1h, NP, C1, C, M;
    const u16x8 *fA, *fC, *iC, *iA;  // [(rt * KS + ks) * 64 + lane][hi, lo]
    const float2 *wN, *rot;          // e^{-2 pi i q/N}, e^{-i pi k/(2N)}
    float scale;                     // inverse: 1/(N M)
};

constexpr int kXL = 2;   // lines per block (even)
constexpr int kLdx = 4;  // loads per thread in flight in the inverse's gather

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulcf(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// split 8 FP32 values into BF16 hi / lo fragments
__device__ __forceinline__ void split8(const float (&x)[8], bf16x8 &hi, bf16x8 &lo) {
    u16x8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const unsigned short b = bf16_bits(x[e]);
        h[e] = b;
        l[e] = bf16_bits(x[e] - bf16_float(b));
    }
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
}

// a b with a = a_hi + a_lo, b = b_hi + b_lo as the three split-BF16 products
// a_lo b_hi + a_hi b_lo + a_hi b_hi (small terms first), FP32 accumulation
__device__ __forceinline__ f32x4 mfma_split(const u16x8 *A, bf16x8 bh, bf16x8 bl, f32x4 acc) {
    const bf16x8 ah = __builtin_bit_cast(bf16x8, A[0]), al = __builtin_bit_cast(bf16x8, A[1]);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// divergence of the first stage's increment U* - Un at cell (i, j, k), p its
// index (the IncSrc form of the pseudo-pressure rhs): c0 F1 on the faces the
// stage update moves, the FP64 increments on the inflow and outflow U faces
// (plus the mass-flux shift on the outflow), the difference of the stored
// values on the wall and top V faces; the W face below plane 1 is the
// periodic image of plane nz-2
__device__ __forceinline__ double inc_div(const Dims &d, const Metrics &mt, const IncSrc &s, long long p, int i,
                                          int j, int k) {
    const long long pl = j + (long long)d.nyg * k, npl = (long long)d.nyg * d.nzg;
    const double ur = (i == d.nx - 2) ? s.inc1[npl + pl] + s.inc1[2 * npl] : s.c0 * (double)s.fu[p];
    const double ul = (i == 1) ? s.inc1[pl] : s.c0 * (double)s.fu[p - 1];
    const double vt = (j == d.ny - 1) ? (double)s.vs[p] - (double)s.vn[p] : s.c0 * (double)s.fv[p];
    const long long pb = p - d.sx;
    const double vb = (j == 1) ? (double)s.vs[pb] - (double)s.vn[pb] : s.c0 * (double)s.fv[pb];
    const long long pw = (k == 1) ? p + d.sxy * (long long)(d.nz - 3) : p - d.sxy;
    const double wt = s.c0 * (double)s.fw[p], wb = s.c0 * (double)s.fw[pw];
    return (ur - ul) * mt.rdx[i] + (vt - vb) * mt.rdy[j] + (wt - wb) * mt.rdz[k];
}

// forward: the pseudo-pressure rhs (projection.f90:38-47) of kXL lines
// (cells i = 1..N of row j, plane m; line c = j-1 + NM (m-1)) formed in FP64
// (from the increment when use_inc), its FP64 line sums for the (0,0) mode
// (part[(j-1) M + m-1]), then the transform into T1[c NP + kx]
__global__ __launch_bounds__(256) void k_xfwd_mx(Dims d, Metrics mt, XArgs a, const float *__restrict__ U,
                                                 const float *__restrict__ V, const float *__restrict__ W,
                                                 float *__restrict__ out, double *__restrict__ part, IncSrc inc,
                                                 int use_inc) {
    extern __shared__ float smx[];
    const int N = a.N, N1 = a.N1, N2 = a.N2, K1 = a.K1h + 1;
    // twiddle indices n2 k1 stay below N (n2 < N2, k1 <= N1/2): no reduction
    const float rN2 = 1.0f / (float)N2, rK1 = 1.0f / (float)K1;
    float *v = smx;                                         // kXL*N
    float2 *Bc = reinterpret_cast<float2 *>(v + kXL * N);   // kXL*K1*N2 (8-byte aligned: kXL*N even)
    __shared__ double lsum[kXL][4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, q = lane >> 4, c16 = lane & 15;
    const int line0 = blockIdx.x * kXL;
    double ps[kXL];
#pragma unroll
    for (int l = 0; l < kXL; ++l) ps[l] = 0.0;
#pragma unroll
    for (int l = 0; l < kXL; ++l) {
        const int c = min(line0 + l, a.C - 1);
        const int c1 = c % a.C1, c2 = c / a.C1;
        const long long base = fidx(d, 1, c1 + 1, c2 + 1);
        const double ry = mt.rdy[c1 + 1], rz = mt.rdz[c2 + 1];
        for (int n = tid; n < N; n += 256) {
            const long long p = base + n;
            const double dv = use_inc ? inc_div(d, mt, inc, p, n + 1, c1 + 1, c2 + 1)
                                      : ((double)U[p] - (double)U[p - 1]) * mt.rdx[n + 1] +
                                            ((double)V[p] - (double)V[p - d.sx]) * ry +
                                            ((double)W[p] - (double)W[p - d.sxy]) * rz;
            ps[l] += dv;
            v[l * N + ((n & 1) ? N - 1 - (n >> 1) : (n >> 1))] = (float)dv;
        }
    }
#pragma unroll
    for (int l = 0; l < kXL; ++l) {
        double t = ps[l];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
        if (lane == 0) lsum[l][wave] = t;
    }
    __syncthreads();
    if (tid < kXL && line0 + tid < a.C) {
        const int c = line0 + tid, c1 = c % a.C1, c2 = c / a.C1;
        part[(long long)c1 * a.M + c2] = (lsum[tid][0] + lsum[tid][1]) + (lsum[tid][2] + lsum[tid][3]);
    }
    // ---- step A: wave w owns column tiles (line, n2) kXL/2 w .. kXL/2 (w+1)
    for (int ct = wave * (kXL / 2); ct < (wave + 1) * (kXL / 2); ++ct) {
        const int col = ct * 16 + c16;
        const int ln = small_div(col, rN2), n2 = col - ln * N2;
        const bool cok = col < kXL * N2;
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int n1 = 8 * q + e;
            x[e] = (cok && n1 < N1) ? v[ln * N + N2 * n1 + n2] : 0.f;
        }
        bf16x8 bh, bl;
        split8(x, bh, bl);
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        const f32x4 acc0 = mfma_split(a.fA + (0 * 64 + lane) * 2, bh, bl, z);
        const f32x4 acc1 = mfma_split(a.fA + (1 * 64 + lane) * 2, bh, bl, z);
        if (cok) {
#pragma unroll
            for (int h = 0; h < 4; ++h) {  // (row tile, pair): k1 = 8 (h/2) + 2 q + h%2
                const f32x4 &acc = (h < 2) ? acc0 : acc1;
                const int k1 = 8 * (h / 2) + 2 * q + (h % 2);
                if (k1 < K1)
                    Bc[(ln * K1 + k1) * N2 + n2] =
                        cmulf(make_float2(acc[2 * (h % 2)], acc[2 * (h % 2) + 1]), a.wN[n2 * k1]);
            }
        }
    }
    __syncthreads();
    // ---- step C: wave w owns row tile w (k2 = 8w + 2q, 8w + 2q + 1), all kXL
    // column tiles (line, k1)
    for (int ct = 0; ct < kXL; ++ct) {
        const int col = ct * 16 + c16;
        const int ln = small_div(col, rK1), k1 = col - ln * K1;
        const bool cok = col < kXL * K1;
        const float *Bl = reinterpret_cast<const float *>(Bc + (ln * K1 + k1) * N2);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            float x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = 32 * ks + 8 * q + e;  // (n2, re/im) = (kk/2, kk%2)
                x[e] = (cok && (kk >> 1) < N2) ? Bl[kk] : 0.f;
            }
            bf16x8 bh, bl;
            split8(x, bh, bl);
            acc = mfma_split(a.fC + ((wave * 2 + ks) * 64 + lane) * 2, bh, bl, acc);
        }
        if (cok) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k2 = 8 * wave + 2 * q + h;
                if (k2 < N2) {
                    const int k = k1 + N1 * k2;
                    const float2 wv = cmulf(a.rot[k], make_float2(acc[2 * h], acc[2 * h + 1]));
                    v[ln * N + k] = 2.f * wv.x;
                    // the mirror N - k of k1 = 0 and k1 = N1/2 is written
                    // directly by another lane (k1' = k1): one writer each
                    if (k1 != 0 && 2 * k1 != N1) v[ln * N + N - k] = -2.f * wv.y;
                }
            }
        }
    }
    __syncthreads();
    for (int idx = tid; idx < kXL * N; idx += 256) {
        const int l = idx >= N ? 1 : 0, kx = idx - l * N;  // kXL = 2
        const int c = line0 + l;
        if (c < a.C) out[(long long)c * a.NP + kx] = v[idx];
    }
}

// inverse: T1 lines (c NP + kx) -> phi cells (1..N, j, m) plus the (0,0)
// constant add[j-1] of each y row
__global__ __launch_bounds__(256) void k_xinv_mx(Dims d, XArgs a, const float *__restrict__ in,
                                                 const double *__restrict__ add, float *__restrict__ phi) {
    extern __shared__ float smx[];
    const int N = a.N, N1 = a.N1, N2 = a.N2, K1h = a.K1h, K1 = K1h + 1;
    // twiddle indices n2 k1 stay below N (n2 < N2, k1 <= N1/2): no reduction
    const float rN2 = 1.0f / (float)N2, rK1 = 1.0f / (float)K1;
    float2 *Vs = reinterpret_cast<float2 *>(smx);  // kXL*K1*N2, later the output lines (kXL*N floats)
    float2 *Cs = Vs + kXL * K1 * N2;               // kXL*K1*N2
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, q = lane >> 4, c16 = lane & 15;
    const int line0 = blockIdx.x * kXL;
    // kLdx elements per thread loaded before the first one is used
    static_assert(kXL == 2, "line index below assumes two lines per block");
    const float *gls[kXL];
#pragma unroll
    for (int l = 0; l < kXL; ++l) gls[l] = in + (long long)min(line0 + l, a.C - 1) * a.NP;
    const int nv = kXL * K1 * N2;
    for (int i0 = tid; i0 < nv; i0 += 256 * kLdx) {
        float gk[kLdx], gm[kLdx];
        float2 rt[kLdx];
#pragma unroll
        for (int u = 0; u < kLdx; ++u) {
            const int idx = min(i0 + 256 * u, nv - 1);
            const int l = idx >= K1 * N2 ? 1 : 0, r = idx - l * K1 * N2;
            const int k1 = small_div(r, rN2), k2 = r - k1 * N2;
            const int k = k1 + N1 * k2;
            gk[u] = gls[l][k];
            gm[u] = gls[l][k == 0 ? 0 : N - k];
            rt[u] = a.rot[k];
            if (k == 0) gm[u] = 0.f;
        }
#pragma unroll
        for (int u = 0; u < kLdx; ++u) {
            const int idx = i0 + 256 * u;
            if (idx < nv) Vs[idx] = cmulcf(make_float2(0.5f * gk[u], -0.5f * gm[u]), rt[u]);
        }
    }
    __syncthreads();
    // ---- step C': wave w owns row tile w (n2 = 8w + 2q, +1), all kXL column
    // tiles (line, k1)
    for (int ct = 0; ct < kXL; ++ct) {
        const int col = ct * 16 + c16;
        const int ln = small_div(col, rK1), k1 = col - ln * K1;
        const bool cok = col < kXL * K1;
        const float *Vl = reinterpret_cast<const float *>(Vs + (ln * K1 + k1) * N2);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            float x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = 32 * ks + 8 * q + e;  // (k2, re/im)
                x[e] = (cok && (kk >> 1) < N2) ? Vl[kk] : 0.f;
            }
            bf16x8 bh, bl;
            split8(x, bh, bl);
            acc = mfma_split(a.iC + ((wave * 2 + ks) * 64 + lane) * 2, bh, bl, acc);
        }
        if (cok) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int n2 = 8 * wave + 2 * q + h;
                if (n2 < N2)
                    Cs[(ln * K1 + k1) * N2 + n2] =
                        cmulcf(make_float2(acc[2 * h], acc[2 * h + 1]), a.wN[n2 * k1]);
            }
        }
    }
    __syncthreads();
    // ---- step A': wave w owns column tiles (line, n2) kXL/2 w .. kXL/2 (w+1),
    // both row tiles (n1 = 16 rt + 4 q + r)
    float *o = smx;  // kXL*N floats inside the Vs area, which is no longer read
    for (int ct = wave * (kXL / 2); ct < (wave + 1) * (kXL / 2); ++ct) {
        const int col = ct * 16 + c16;
        const int ln = small_div(col, rN2), n2 = col - ln * N2;
        const bool cok = col < kXL * N2;
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int kk = 8 * q + e, k1 = kk >> 1;  // (k1, re/im)
            x[e] = (cok && k1 < K1) ? reinterpret_cast<const float *>(Cs + (ln * K1 + k1) * N2 + n2)[kk & 1] : 0.f;
        }
        bf16x8 bh, bl;
        split8(x, bh, bl);
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        const f32x4 acc0 = mfma_split(a.iA + (0 * 64 + lane) * 2, bh, bl, z);
        const f32x4 acc1 = mfma_split(a.iA + (1 * 64 + lane) * 2, bh, bl, z);
        if (cok) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int n1 = 16 * (r / 4) + 4 * q + (r % 4);
                if (n1 < N1) {
                    const int n = N2 * n1 + n2;
                    const int pos = (n < (N + 1) / 2) ? 2 * n : 2 * (N - 1 - n) + 1;
                    o[ln * N + pos] = (r < 4 ? acc0[r % 4] : acc1[r % 4]) * a.scale;
                }
            }
        }
    }
    __syncthreads();
    // per-line output bases and (0,0) constants, then one element per thread
    // and step (kXL = 2)
    long long ob[kXL];
    float ad[kXL];
#pragma unroll
    for (int l = 0; l < kXL; ++l) {
        const int c = min(line0 + l, a.C - 1), c1 = c % a.C1, c2 = c / a.C1;
        ob[l] = fidx(d, 1, c1 + 1, c2 + 1);
        ad[l] = (float)add[c1];
    }
    for (int idx = tid; idx < kXL * N; idx += 256) {
        const int l = idx >= N ? 1 : 0, n = idx - l * N;
        if (line0 + l >= a.C) continue;
        phi[ob[l] + n] = o[idx] + ad[l];
    }
}

// ------------------------------------------------------------------------
// pseudo-pressure rhs (projection.f90:38-47) in compact lines
// DIV[(m*NM + j)*NP + i] (cells i+1, j+1, m+1), and per x-z plane FP64
// partial sums of the same divergence for the (0,0) mode: one wave = one y
// row of 64 cells, part[j][bx + gx*m].
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_div_mixed(Dims d, Metrics m, const float *__restrict__ U,
                                                   const float *__restrict__ V, const float *__restrict__ W,
                                                   int NP, float *__restrict__ div, double *__restrict__ part,
                                                   IncSrc inc, int use_inc) {
    const int ii = blockIdx.x * 64 + (threadIdx.x & 63);
    const int jj = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int mm = blockIdx.z;
    double v = 0.0;
    if (ii < d.N && jj < d.NM) {
        const int i = ii + 1, j = jj + 1, k = mm + 1;
        const long long p = fidx(d, i, j, k);
        v = use_inc ? inc_div(d, m, inc, p, i, j, k)
                    : ((double)U[p] - (double)U[p - 1]) * m.rdx[i] + ((double)V[p] - (double)V[p - d.sx]) * m.rdy[j] +
                          ((double)W[p] - (double)W[p - d.sxy]) * m.rdz[k];
        div[((long long)mm * d.NM + jj) * NP + ii] = (float)v;
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if ((threadIdx.x & 63) == 0 && jj < d.NM)
        part[(long long)jj * gridDim.x * gridDim.z + blockIdx.x + (long long)gridDim.x * mm] = v;
}

// FP64 (0,0) mode, in two kernels.  k_sum00: rhs(j) = 2 sum over the x-z
// plane (DCT-II x2 at kx = 0, DFT at kz = 0) of the divergence partials, one
// block per y row, fixed summation order.  Then the pinned kx = kz = 0
// system (launch_solve00) turns it into the constant phi00(j) / (2 N M) that
// the inverse transforms would add.
__global__ __launch_bounds__(256) void k_sum00(const double *__restrict__ part, int np,
                                               double *__restrict__ r00) {
    __shared__ double w[4];
    const int j = blockIdx.x;
    const double *p = part + (long long)j * np;
    double s = 0.0;
    for (int q = threadIdx.x; q < np; q += 256) s += p[q];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) r00[j] = 2.0 * ((w[0] + w[1]) + (w[2] + w[3]));
}

// Thomas sweep in FP64 on the FP32 spectrum (kx stride NP): the forward
// values are kept in FP64 in X ([q][j][kx], kx stride N) for the back
// substitution.  The (0,0) system is zeroed here (launch_solve00 solves it).
template <int TJ>
__global__ __launch_bounds__(256) void k_thomas_mixed(Dims d, int NP, const double *__restrict__ dl,
                                                      const double *__restrict__ dp,
                                                      const double *__restrict__ du, float *S,
                                                      double *__restrict__ X) {
    const int N = d.N, NM = d.NM;
    int kxb, q;
    thomas_block(blockIdx.x, (N + 255) / 256, d.M, kxb, q);
    const int kx = kxb * 256 + threadIdx.x;
    if (kx >= N) return;
    const long long sp = (long long)NP * NM, fp = (long long)N * NM;
    float *a = S + q * sp + kx;
    if (kx == 0 && q == 0) {
        for (int j = 0; j < NM; ++j) a[(long long)j * NP] = 0.f;
        return;
    }
    const int kk = (q + 1) / 2;
    const double *L = dl + kk * fp + kx;
    const double *D = dp + kk * fp + kx;
    double *x = X + q * fp + kx;
    double r = a[0];
    x[0] = r;
    int j = 1;
    for (; j + TJ <= NM; j += TJ) {
        double v[TJ], l[TJ];
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            v[t] = a[(long long)(j + t) * NP];
            l[t] = L[(long long)(j + t - 1) * N];
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            r = v[t] - l[t] * r;
            x[(long long)(j + t) * N] = r;
        }
    }
    for (; j < NM; ++j) {
        r = (double)a[(long long)j * NP] - L[(long long)(j - 1) * N] * r;
        x[(long long)j * N] = r;
    }
    r = r * D[(long long)(NM - 1) * N];
    a[(long long)(NM - 1) * NP] = (float)r;
    j = NM - 2;
    for (; j - TJ + 1 >= 0; j -= TJ) {
        double v[TJ], p[TJ], u[TJ];
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            v[t] = x[(long long)(j - t) * N];
            p[t] = D[(long long)(j - t) * N];
            u[t] = du[j - t];
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            r = (v[t] - u[t] * r) * p[t];
            a[(long long)(j - t) * NP] = (float)r;
        }
    }
    for (; j >= 0; --j) {
        r = (x[(long long)j * N] - du[j] * r) * D[(long long)j * N];
        a[(long long)j * NP] = (float)r;
    }
}

void check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        last_error() = std::string("kernel launch: ") + hipGetErrorString(e);
        throw DeviceError();
    }
}

// host: BF16 hi/lo bits of a float (round to nearest even; finite inputs)
uint16_t host_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float host_bf16_float(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// rows x K matrix (row-major, entries from f(r, k)) into hi/lo BF16 rows of
// length Kp, rows padded to Rp; padding is zero
template <class F>
void split_matrix(int rows, int K, int Rp, int Kp, F f, std::vector<uint16_t> &hi, std::vector<uint16_t> &lo) {
    hi.assign((size_t)Rp * Kp, 0);
    lo.assign((size_t)Rp * Kp, 0);
    for (int r = 0; r < rows; ++r)
        for (int k = 0; k < K; ++k) {
            const float v = (float)f(r, k);
            const uint16_t h = host_bf16(v);
            hi[(size_t)r * Kp + k] = h;
            lo[(size_t)r * Kp + k] = host_bf16(v - host_bf16_float(h));
        }
}

// x: 128 x 128 tiles (812 rows -> 7 row tiles), K step 32 (two waves per
// SIMD); z: 256 x 128 tiles, all 191 rows in one row tile so each B column
// block is read once, K step 64 (110.6 KB of LDS, one block per CU)
constexpr int kWRx = 2, kWCx = 2, kBKx = 32, kWRz = 4, kWCz = 2, kBKz = 64;
constexpr int kBMx = 64 * kWRx, kBMz = 64 * kWRz;

}  // namespace

void build_mixed(Handle &h) {
    const Dims &d = h.d;
    const int N = d.N, M = d.M;
    if (d.NM > kMax00) {
        last_error() = "bl_create: the mixed-precision path supports ny <= 1025";
        throw ArgError();
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    h.NP = (N + 31) / 32 * 32;
    h.kpx = (N + kBKx - 1) / kBKx * kBKx;
    h.kpz = (M + kBKz - 1) / kBKz * kBKz;
    const int rpx = (N + kBMx - 1) / kBMx * kBMx, rpz = (M + kBMz - 1) / kBMz * kBMz;
    // exact integer argument reduction before the long double cosine
    auto cosx = [&](int k, int n) {  // cos(pi k (2n+1) / (2N))
        const long long a = ((long long)k * (2 * n + 1)) % (4LL * N);
        return cosl(pi * (long double)a / (2.0L * N));
    };
    auto cz = [&](int k, int m) { return cosl(2.0L * pi * (long double)(((long long)k * m) % M) / M); };
    auto sz = [&](int k, int m) { return sinl(2.0L * pi * (long double)(((long long)k * m) % M) / M); };
    const long double nm = (long double)N * (long double)M;
    std::vector<uint16_t> t[8];
    split_matrix(N, N, rpx, h.kpx, [&](int k, int n) { return 2.0L * cosx(k, n); }, t[0], t[1]);
    split_matrix(N, N, rpx, h.kpx,
                 [&](int n, int k) { return (k == 0 ? 0.5L : 1.0L) * cosx(k, n) / nm; }, t[2], t[3]);
    split_matrix(M, M, rpz, h.kpz,
                 [&](int q, int m) -> long double {
                     if (q == 0) return 1.0L;
                     const int k = (q + 1) / 2;
                     return (q % 2 == 1) ? cz(k, m) : -sz(k, m);
                 },
                 t[4], t[5]);
    split_matrix(M, M, rpz, h.kpz,
                 [&](int m, int q) -> long double {
                     if (q == 0) return 1.0L;
                     const int k = (q + 1) / 2;
                     if (q % 2 == 1) return (2 * k == M ? 1.0L : 2.0L) * cz(k, m);
                     return -2.0L * sz(k, m);
                 },
                 t[6], t[7]);
    size_t total = 0;
    for (auto &v : t) total += v.size();
    BL_HIP(hipMalloc(&h.mx_tab, total * sizeof(uint16_t)));
    const unsigned short **dst[8] = {&h.dxf_hi, &h.dxf_lo, &h.dxi_hi, &h.dxi_lo,
                                     &h.zf_hi,  &h.zf_lo,  &h.zi_hi,  &h.zi_lo};
    size_t off = 0;
    for (int q = 0; q < 8; ++q) {
        BL_HIP(hipMemcpyAsync(h.mx_tab + off, t[q].data(), t[q].size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                              h.stream));
        *dst[q] = h.mx_tab + off;
        off += t[q].size();
    }
    // folded z transform fragments (k_zdft_mx) for M/2 + 1 <= 96: per (row
    // tile t, k step s, lane l) the 8 values of A[16 t + l % 16][32 s + 8 (l / 16) + e]
    // as cos hi, cos lo, sin hi, sin lo (zero past M/2)
    if (M / 2 + 1 <= 16 * kZxRT) {
        const int Kc = M / 2 + 1, RT = (Kc + 15) / 16, KS = (Kc + 31) / 32;
        std::vector<uint16_t> zf((size_t)RT * KS * 64 * 32, 0);
        for (int tt = 0; tt < RT; ++tt)
            for (int ss = 0; ss < KS; ++ss)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int row = 16 * tt + l % 16, kk = 32 * ss + 8 * (l / 16) + e;
                        if (kk >= Kc) continue;
                        const size_t o = (((size_t)tt * KS + ss) * 64 + l) * 32 + e;
                        const float vc = (float)cz(row, kk), vs = (float)sz(row, kk);
                        const uint16_t ch = host_bf16(vc), sh = host_bf16(vs);
                        zf[o] = ch;
                        zf[o + 8] = host_bf16(vc - host_bf16_float(ch));
                        zf[o + 16] = sh;
                        zf[o + 24] = host_bf16(vs - host_bf16_float(sh));
                    }
        BL_HIP(hipMalloc(&h.zmf, zf.size() * sizeof(uint16_t)));
        BL_HIP(hipMemcpyAsync(h.zmf, zf.data(), zf.size() * sizeof(uint16_t), hipMemcpyHostToDevice, h.stream));
        h.zmf_rt = RT;
    }
    // two-factor x transform (k_xfwd_mx / k_xinv_mx) when N = N1 N2 fits it
    {
        // N = N1 N2 with N1 the largest divisor <= sqrt(N), as for the FP64 tables
        int n1 = 1;
        for (int a = 1; (long long)a * a <= N; ++a)
            if (N % a == 0) n1 = a;
        h.dct_n1 = n1;
        const int n2 = N / n1, k1n = n1 / 2 + 1;
        h.mx_two = n1 <= 32 && k1n <= 16 && n2 <= 32;
        if (h.mx_two) {
            auto cis = [&](long long num, long long den) {  // e^{-2 pi i num/den}, exact reduction
                const long double ang = -2.0L * pi * (long double)(num % den) / (long double)den;
                return std::pair<long double, long double>(cosl(ang), sinl(ang));
            };
            // one table: tiles x ks x 64 lanes x (hi, lo) x 8 k values
            auto frag = [&](int tiles, int ks_n, auto f) {
                std::vector<uint16_t> t((size_t)tiles * ks_n * 64 * 16, 0);
                for (int rt = 0; rt < tiles; ++rt)
                    for (int ks = 0; ks < ks_n; ++ks)
                        for (int l = 0; l < 64; ++l)
                            for (int e = 0; e < 8; ++e) {
                                const float v = (float)f(rt, l % 16, 32 * ks + 8 * (l / 16) + e);
                                const uint16_t hb = host_bf16(v);
                                const size_t o = (((size_t)rt * ks_n + ks) * 64 + l) * 16;
                                t[o + e] = hb;
                                t[o + 8 + e] = host_bf16(v - host_bf16_float(hb));
                            }
                return t;
            };
            auto pair_row = [](int rt, int R, int &idx, int &part) {
                idx = 8 * rt + 2 * (R / 4) + (R % 4) / 2;
                part = R % 2;
            };
            std::vector<uint16_t> fA = frag(2, 1, [&](int rt, int R, int k) -> long double {
                int k1, part;
                pair_row(rt, R, k1, part);
                if (k1 >= k1n || k >= n1) return 0.0L;
                const auto w = cis((long long)k * k1, n1);
                return part == 0 ? w.first : w.second;
            });
            std::vector<uint16_t> fC = frag(4, 2, [&](int rt, int R, int k) -> long double {
                int k2, part;
                pair_row(rt, R, k2, part);
                const int nn = k / 2, pin = k % 2;
                if (k2 >= n2 || nn >= n2) return 0.0L;
                const auto w = cis((long long)nn * k2, n2);
                return part == 0 ? (pin == 0 ? w.first : -w.second) : (pin == 0 ? w.second : w.first);
            });
            std::vector<uint16_t> iC = frag(4, 2, [&](int rt, int R, int k) -> long double {
                int nn, part;
                pair_row(rt, R, nn, part);
                const int k2 = k / 2, pin = k % 2;
                if (nn >= n2 || k2 >= n2) return 0.0L;
                const auto w = cis((long long)nn * k2, n2);
                return part == 0 ? (pin == 0 ? w.first : w.second) : (pin == 0 ? -w.second : w.first);
            });
            std::vector<uint16_t> iA = frag(2, 1, [&](int rt, int R, int k) -> long double {
                const int nn1 = 16 * rt + R, k1 = k / 2, pin = k % 2;
                if (nn1 >= n1 || k1 >= k1n) return 0.0L;
                const long double c = (k1 == 0 || (n1 % 2 == 0 && k1 == n1 / 2)) ? 1.0L : 2.0L;
                const auto w = cis((long long)nn1 * k1, n1);
                return c * (pin == 0 ? w.first : w.second);
            });
            std::vector<float> tw(4 * (size_t)N);
            for (int q = 0; q < N; ++q) {
                const auto w = cis(q, N), r = cis(q, 4LL * N);
                tw[2 * q] = (float)w.first;
                tw[2 * q + 1] = (float)w.second;
                tw[2 * N + 2 * q] = (float)r.first;
                tw[2 * N + 2 * q + 1] = (float)r.second;
            }
            const size_t nb = (fA.size() + fC.size() + iC.size() + iA.size()) * sizeof(uint16_t) + tw.size() * 4;
            BL_HIP(hipMalloc(&h.mx2, nb));
            char *base = reinterpret_cast<char *>(h.mx2);
            size_t off = 0;
            const void **dst[4] = {(const void **)&h.x2_fA, (const void **)&h.x2_fC, (const void **)&h.x2_iC,
                                   (const void **)&h.x2_iA};
            std::vector<uint16_t> *src[4] = {&fA, &fC, &iC, &iA};
            for (int t = 0; t < 4; ++t) {
                BL_HIP(hipMemcpyAsync(base + off, src[t]->data(), src[t]->size() * 2, hipMemcpyHostToDevice, h.stream));
                *dst[t] = base + off;
                off += src[t]->size() * 2;
            }
            BL_HIP(hipMemcpyAsync(base + off, tw.data(), tw.size() * 4, hipMemcpyHostToDevice, h.stream));
            h.x2_tw = reinterpret_cast<const float *>(base + off);
        }
    }
    // FP64 (0,0) mode scratch: per-line (two-factor path) or per-block plane
    // partials of the divergence, then phi00/(2NM)
    h.np00 = h.mx_two ? M : ((N + 63) / 64) * M;
    BL_HIP(hipMalloc(&h.p00, sizeof(double) * ((size_t)h.np00 * d.NM + d.NM)));
    h.c00 = h.p00 + (size_t)h.np00 * d.NM;
    // first-stage boundary increments (written by the boundary update of
    // every first stage before they are read)
    const size_t ninc = 2 * (size_t)d.nyg * d.nzg + 1;
    BL_HIP(hipMalloc(&h.inc1, sizeof(double) * ninc));
    BL_HIP(hipMemsetAsync(h.inc1, 0, sizeof(double) * ninc, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
}

void launch_poisson_mixed(Handle &h, float *phi, const Fld3<float> &f, const IncSrc *inc) {
    const Dims &d = h.d;
    const IncSrc src = inc ? *inc : IncSrc{};
    const int use_inc = inc ? 1 : 0;
    const int N = d.N, M = d.M, NM = d.NM, NP = h.NP;
    float *S = reinterpret_cast<float *>(h.S);
    float *T1 = reinterpret_cast<float *>(h.T1);
    XArgs xa{};
    if (h.mx_two) {
        xa.N = N;
        xa.N1 = h.dct_n1;
        xa.N2 = N / h.dct_n1;
        xa.K1h = h.dct_n1 / 2;
        xa.NP = NP;
        xa.C1 = NM;
        xa.C = NM * M;
        xa.M = M;
        xa.fA = reinterpret_cast<const u16x8 *>(h.x2_fA);
        xa.fC = reinterpret_cast<const u16x8 *>(h.x2_fC);
        xa.iC = reinterpret_cast<const u16x8 *>(h.x2_iC);
        xa.iA = reinterpret_cast<const u16x8 *>(h.x2_iA);
        xa.wN = reinterpret_cast<const float2 *>(h.x2_tw);
        xa.rot = reinterpret_cast<const float2 *>(h.x2_tw) + N;
        xa.scale = (float)(1.0 / ((double)N * (double)M));
    }
    const int K1 = h.dct_n1 / 2 + 1, N2 = N / h.dct_n1;
    const dim3 xgrid((NM * M + kXL - 1) / kXL);
    if (h.mx_two) {
        // divergence, FP64 line sums and forward x transform in one kernel
        ProfScope ps(h, KID_XFORM_X_FWD);
        const size_t lds = sizeof(float) * (size_t)kXL * N + sizeof(float2) * (size_t)kXL * K1 * N2;
        hipLaunchKernelGGL(k_xfwd_mx, xgrid, dim3(256), lds, h.stream, d, h.m, xa, f.u, f.v, f.w, T1, h.p00, src,
                           use_inc);
        check();
    }
    // rhs: divergence lines into S (as scratch), FP64 plane partials
    if (!h.mx_two) {
        ProfScope ps(h, KID_DIVERGENCE);
        hipLaunchKernelGGL(k_div_mixed, dim3((N + 63) / 64, (NM + 3) / 4, M), dim3(256), 0, h.stream, d, h.m,
                           f.u, f.v, f.w, NP, S, h.p00, src, use_inc);
        check();
    }
    // forward x: T1[(m*NM + j)*NP + kx] = sum_n Dxf[kx][n] div[(m*NM + j)*NP + n]
    GemmArgs g{};
    g.R = N;
    g.C = NM * M;
    g.K = N;
    g.Kp = h.kpx;
    g.Ahi = h.dxf_hi;
    g.Alo = h.dxf_lo;
    g.B = S;
    g.ldb = NP;
    g.out = T1;
    g.o_base = 0;
    g.o_c1 = NP;
    g.o_c2 = 0;
    g.o_C1 = g.C;
    g.add = nullptr;
    if (!h.mx_two) gemm<kWRx, kWCx, kBKx, true, true>(h, KID_XFORM_X_FWD, g);
    // forward z: S[q][col] = sum_m Zf[q][m] T1[m][col], col = j*NP + kx
    GemmArgs z{};
    z.R = M;
    z.C = NM * NP;
    z.K = M;
    z.Kp = h.kpz;
    z.Ahi = h.zf_hi;
    z.Alo = h.zf_lo;
    z.B = T1;
    z.ldb = (long long)NM * NP;
    z.out = S;
    z.ldc = (long long)NM * NP;
    z.o_C1 = 1;
    if (h.zmf)
        zdft_mx_launch<kZxRT>(h, KID_XFORM_Z_FWD, h.zmf_rt, M, z.C, z.ldb, true, T1, S);
    else
        gemm<kWRz, kWCz, kBKz, false, false>(h, KID_XFORM_Z_FWD, z);
    // FP64 Thomas on the spectrum, FP64 (0,0) mode
    {
        ProfScope ps(h, KID_THOMAS);
        const double s00 = 1.0 / (2.0 * (double)N * (double)M);
        if (!launch_thomas_seg<float>(h, S, NP, (long long)NP * NM, solve00_args(h, h.c00, 1, s00, h.p00, h.np00))) {
            hipLaunchKernelGGL(k_thomas_mixed<8>, dim3((N + 255) / 256 * M), dim3(256), 0, h.stream, d, NP,
                               h.lu_dl, h.lu_dp, h.lu_du, S, h.T1);
            check();
            hipLaunchKernelGGL(k_sum00, dim3(NM), dim3(256), 0, h.stream, h.p00, h.np00, h.c00);
            check();
            launch_solve00(h, solve00_args(h, h.c00, 1, s00, nullptr, 0));
        }
    }
    // inverse z: T1[m][col] = sum_q Zi[m][q] S[q][col]
    z.Ahi = h.zi_hi;
    z.Alo = h.zi_lo;
    z.B = S;
    z.out = T1;
    if (h.zmf)
        zdft_mx_launch<kZxRT>(h, KID_XFORM_Z_INV, h.zmf_rt, M, z.C, z.ldb, false, S, T1);
    else
        gemm<kWRz, kWCz, kBKz, false, false>(h, KID_XFORM_Z_INV, z);
    // inverse x into the field layout, plus the (0,0) constant of each y row
    g.Ahi = h.dxi_hi;
    g.Alo = h.dxi_lo;
    g.B = T1;
    g.out = phi;
    g.o_base = fidx(d, 1, 1, 1);
    g.o_C1 = NM;
    g.o_c1 = d.sx;
    g.o_c2 = d.sxy;
    g.add = h.c00;
    if (h.mx_two) {
        ProfScope ps(h, KID_XFORM_X_INV);
        const size_t lds = 2 * sizeof(float2) * (size_t)kXL * K1 * N2;
        hipLaunchKernelGGL(k_xinv_mx, xgrid, dim3(256), lds, h.stream, d, xa, T1, h.c00, phi);
        check();
    } else {
        gemm<kWRx, kWCx, kBKx, true, true>(h, KID_XFORM_X_INV, g);
    }
}

}  // namespace blgpu
