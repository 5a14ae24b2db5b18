// kern_struct.hpp -- S.x and O statistics of the basic RBM without the O
// matrix, as bf16 MFMA GEMMs (gfx950 v_mfma_f32_16x16x32_bf16) at FP32
// accuracy.
//
// For the basic RBM (rbm_base.cpp:194-217) the O column of sample k is
//   O_k = [ sigma_k (N) | t_k (M) | vec(sigma_k t_k^T) (N x M, j + i N) ],
// t_k = tanh theta_k, so with x = [x_v | x_h | X (N x M)]:
//   y_k = O_k . x = sigma_k . x_v + sum_i t_ki (x_hi + (Sigma X)_ki)
//   sum_k conj(O_k) y_k = [ Sigma^T y | conj(T)^T y | Sigma^T diag(y) conj(T) ].
// Sigma (S x N, entries +-1) and X / T are the only operands: one CG matvec
// reads the sigma bits and T (S x M complex) once instead of O (S x P).
//
// Precision: sigma = +-1 is exact in bf16.  The other operand of each GEMM
// (X in stage 1, y conj(t) in stage 2) is split by truncation into three
// bf16 terms hi + mid + lo, which hold the 24 significant bits of an FP32
// value exactly; every product is then exact and the MFMA accumulates in
// FP32, so the result carries FP32 rounding only (three bf16 MFMAs cost
// less than one FP32 MFMA of the same size).
//
// Work split: each wave walks chunks of 32 samples (two tiles of 16) on
// its own; no LDS access or barrier inside the loop.
//  * stage 1 (mode 0): G = Sigma_tile X.  A = the tile's sigma rows, built
//    from the state bits in registers; B = the X terms, staged once in LDS.
//    Accumulator register r of lane l holds G[4(l/16)+r][column l%16].
//  * the lane loads exactly the T entries of that layout (one chunk ahead
//    for the smaller shapes), forms sum_i t (x_h + G) + sigma . x_v over
//    its columns, and a 16-lane reduction gives y for samples 4(l/16)+r of
//    both tiles.
//  * stage 2: Z += Sigma^T diag(y) conj(T) with the chunk's 32 samples as
//    the K dimension.  Lane l supplies K slots (l/16, j), j < 8, which are
//    assigned the samples 16(j/4) + 4(l/16) + j%4 -- the ones whose t and
//    y the lane already holds -- so B = y conj(t) is formed in registers.
//    (A and B use the same slot -> sample assignment; the order of the K
//    dimension does not change the product.)
//  * column sums Sigma^T y, conj(T)^T y and sum |t|^2 accumulate in VALU
//    registers; the waves of a block (one block per CU) add their partials
//    in a fixed tree through LDS and write one partial row per block.
// Hidden-unit order: column n of column tile c is hidden unit
// 32(c/2) + 2n + c%2 (CT even), so one 16-byte load fetches a lane's
// entries of tiles c = 2p and 2p + 1.
// Modes: 0 = S.x (y from x), 1 = y = E_loc (gives sum E conj(O)),
// 2 = y = 1 (gives sum conj(O) and sum |t|^2).
//
// MFMA operand layout (16x16x32 bf16): lane l supplies A[l%16][8(l/16)+j]
// and B[8(l/16)+j][l%16], j < 8 (element j of the 8-element operand);
// accumulator register r of lane l is D[4(l/16)+r][l%16].
#pragma once
#include "dev_common.hpp"
#include "kern_api.hpp"

namespace rbmgpu {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// fourth word of a raw buffer resource on gfx9 (32-bit data format; the
// range check makes loads past num_records return 0)
constexpr int kBufWord3 = 0x00020000;

__device__ inline v4f mfma_bf16(v4u a, v4u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                   0);
}

// bf16 truncations of a (low half) and b (high half), packed
__device__ inline unsigned pack_hi(float a, float b) {
    return __builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), 0x07060302u);
}
__device__ inline float trunc_bf16(float a) { return __uint_as_float(__float_as_uint(a) & 0xffff0000u); }

// (a, b) = hi + mid + lo exactly, each a packed pair of bf16: a - trunc(a)
// is exact (same exponent, low mantissa bits) and keeps at most 16
// significant bits, its remainder at most 8, which bf16 holds exactly
struct Split3 {
    unsigned hi, mid, lo;
};
__device__ inline Split3 split3(float a, float b) {
    Split3 s;
    s.hi = pack_hi(a, b);
    a -= trunc_bf16(a);
    b -= trunc_bf16(b);
    s.mid = pack_hi(a, b);
    a -= trunc_bf16(a);
    b -= trunc_bf16(b);
    s.lo = pack_hi(a, b);
    return s;
}

// two bf16 +-1 from sigma bits b0 (low half), b1 (high half); bit set = -1
__device__ inline unsigned pm1_pair(unsigned b0, unsigned b1) { return 0x3F803F80u | (b0 << 15) | (b1 << 31); }

// stage-2 MFMA over the 16 TA samples of a chunk: K = 32 (16x16x32, eight
// bf16 per lane) for two tiles, K = 16 (16x16x16, four per lane) for one
template <int TA>
struct StageTwo;
template <>
struct StageTwo<2> {
    typedef v4u V;
    __device__ static v4f mfma(V a, V b, v4f c) { return mfma_bf16(a, b, c); }
};
template <>
struct StageTwo<1> {
    typedef v2u V;
    typedef short v4s __attribute__((ext_vector_type(4)));
    __device__ static v4f mfma(V a, V b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(v4s, a), __builtin_bit_cast(v4s, b), c,
                                                          0, 0, 0);
    }
};

template <int CT, int RT, int TA = 1>
struct StructCfg {
    static constexpr int Mp = 16 * CT, Np = 16 * RT;
    static constexpr int KC = (Np + 31) / 32;  // stage-1 K chunks of 32 spins
    static constexpr int CP = CT % 2 == 0 ? 2 : 1;
    // per-lane partial floats: Z (re, im), z_v, z_h, sum |t|^2
    static constexpr int oZV = RT * CT * 8, oZH = oZV + 2 * RT, oT2 = oZH + 2 * CT, NQ = oT2 + CT;
    static constexpr int RS = 65;  // LDS stride of one partial float over the 64 lanes
    // waves per SIMD the registers are budgeted for: two (256 registers)
    // hide each other's latencies; larger shapes would spill at two.  A
    // block holds the waves of one CU (4 SIMDs), so its partial row covers
    // as many samples as possible.
    static constexpr int WPE = RT * CT <= 8 ? 2 : 1;
    static constexpr int NWB = 4 * WPE, THREADS = 64 * NWB;
    static constexpr size_t xbytes = (size_t)KC * CT * 6 * 64 * 16;
    // tree reduction over the block's waves: NWB / 2 regions of NQ floats x 64 lanes
    static constexpr size_t rbytes = (size_t)(NWB / 2) * NQ * RS * 4;
    static constexpr size_t bytes = xbytes > rbytes ? xbytes : rbytes;
    // hidden unit of column n of column tile c, and its inverse
    __host__ __device__ static constexpr int hid(int c, int n) { return 16 * CP * (c / CP) + CP * n + c % CP; }
    __device__ static void inv(int i, int& c, int& n) {
        const int p = i / (16 * CP), rem = i % (16 * CP);
        n = rem / CP;
        c = CP * p + rem % CP;
    }
};

template <int CT, int RT, int TA>
__global__ void __launch_bounds__((StructCfg<CT, RT, TA>::THREADS), (StructCfg<CT, RT, TA>::WPE))
    k_struct(StructArgs a) {
    using C = StructCfg<CT, RT, TA>;
    constexpr int NWB = C::NWB, NT = C::THREADS;
    constexpr int kChunk = 16 * TA;  // samples per wave step
    constexpr int KC = C::KC, CP = C::CP;
    static_assert(RT <= 4, "one state word per sample (N <= 64)");
    extern __shared__ __align__(16) v4u xs[];  // [KC][CT][6][64] X terms, then the block reduction
    if (a.status && a.status->done) return;
    const int N = a.N, M = a.M, mode = a.mode, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = lane & 15, g = lane >> 4;
    // ---- X terms: B operand of stage 1 for K chunk kc and column tile c;
    // lane ln supplies X[32 kc + 8 (ln/16) + j][hid(c, ln%16)], j < 8, as
    // re hi, mid, lo, im hi, mid, lo
    if (mode == 0) {
        for (int e = tid; e < KC * CT * 64; e += NT) {
            const int ln = e & 63, c = (e >> 6) % CT, kc = (e >> 6) / CT;
            const int i = C::hid(c, ln & 15), j0 = 32 * kc + 8 * (ln >> 4);
            // all 16 loads issued unconditionally (an index past the block
            // reads entry 0), then masked
            double2 xq[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int sp = j0 + j;
                const int64_t q = (sp < N && i < M) ? (int64_t)N + M + sp + (int64_t)i * N : 0;
                xq[j] = reinterpret_cast<const double2*>(a.x)[q];
            }
            float re[8], im[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const bool in = j0 + j < N && i < M;
                re[j] = in ? (float)xq[j].x : 0.f;
                im[j] = in ? (float)xq[j].y : 0.f;
            }
            v4u t[6];
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const Split3 sr = split3(re[2 * h], re[2 * h + 1]), si = split3(im[2 * h], im[2 * h + 1]);
                t[0][h] = sr.hi;
                t[1][h] = sr.mid;
                t[2][h] = sr.lo;
                t[3][h] = si.hi;
                t[4][h] = si.mid;
                t[5][h] = si.lo;
            }
#pragma unroll
            for (int q = 0; q < 6; q++) xs[((kc * CT + c) * 6 + q) * 64 + ln] = t[q];
        }
    }
    // x_h of the lane's hidden units, x_v of its spins n + 16 q
    cplx<float> xh[CT], xv[RT];
    if (mode == 0) {
        const double2* x2 = reinterpret_cast<const double2*>(a.x);
        double2 lh[CT], lv[RT];
#pragma unroll
        for (int c = 0; c < CT; c++) lh[c] = x2[N + min(C::hid(c, n), M - 1)];
#pragma unroll
        for (int q = 0; q < RT; q++) lv[q] = x2[min(n + 16 * q, N - 1)];
#pragma unroll
        for (int c = 0; c < CT; c++)
            xh[c] = C::hid(c, n) < M ? cplx<float>{(float)lh[c].x, (float)lh[c].y} : cplx<float>{0.f, 0.f};
#pragma unroll
        for (int q = 0; q < RT; q++)
            xv[q] = n + 16 * q < N ? cplx<float>{(float)lv[q].x, (float)lv[q].y} : cplx<float>{0.f, 0.f};
    } else {
#pragma unroll
        for (int c = 0; c < CT; c++) xh[c] = {0.f, 0.f};
#pragma unroll
        for (int q = 0; q < RT; q++) xv[q] = {0.f, 0.f};
    }
    __syncthreads();

    v4f zr[RT][CT], zi[RT][CT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int c = 0; c < CT; c++) zr[rt][c] = zi[rt][c] = v4f{0, 0, 0, 0};
    cplx<float> zh[CT], zv[RT];
    float t2[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        zh[c] = {0.f, 0.f};
        t2[c] = 0.f;
    }
#pragma unroll
    for (int q = 0; q < RT; q++) zv[q] = {0.f, 0.f};

    // one chunk's operands: t of samples 16 ta + 4 g + r (all columns of
    // the lane), their state words, and the words of samples 16 ta + n.
    // Raw buffer loads: the chunk base and the row offsets 16 ta + r are
    // wave-uniform (scalar resource and offset), the lane offsets are fixed
    // for the whole kernel, and no load carries a condition (a condition
    // puts each load behind a branch whose join waits for every load in
    // flight).  Rows past S lie outside the resource's range and read as 0;
    // those samples also get y = 0.  A padded hidden unit (hid >= M) reads
    // unit M - CP instead; it meets X = 0 and x_h = 0 in stage 1 and its Z,
    // z_h and sum |t|^2 entries are not written.
    struct Chunk {
        cplx<float> t[TA][4][CT];
        uint64_t ws[TA][4];
        uint64_t wa[TA];
    };
    // byte offset of the lane's units in row 4 g (+ 16 ta + r) of the chunk
    uint32_t toff[CT / CP];
#pragma unroll
    for (int c = 0; c < CT; c += CP) toff[c / CP] = (uint32_t)(4 * g * M + min(C::hid(c, n), M - CP)) * 8u;
    auto fetch = [&](int64_t ch, Chunk& d) {
        const int64_t s0 = ch * kChunk;
        // rows of the chunk inside [0, S): 0 for a chunk past the end, whose
        // loads then return 0 without a memory access
        const int rows = (int)max(min(a.S - s0, (int64_t)kChunk), (int64_t)0);
        const __amdgpu_buffer_rsrc_t rT = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<cplx<float>*>(a.T + s0 * M), 0, rows * M * 8, kBufWord3);
        const __amdgpu_buffer_rsrc_t rS =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(a.states + s0), 0, rows * 8, kBufWord3);
#pragma unroll
        for (int ta = 0; ta < TA; ta++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int so = (16 * ta + r) * M * 8;
#pragma unroll
                for (int c = 0; c < CT; c += CP) {
                    if constexpr (CP == 2) {
                        const v4u v = __builtin_amdgcn_raw_buffer_load_b128(rT, toff[c / 2], so, 0);
                        d.t[ta][r][c] = {__uint_as_float(v.x), __uint_as_float(v.y)};
                        d.t[ta][r][c + 1] = {__uint_as_float(v.z), __uint_as_float(v.w)};
                    } else {
                        const v2u v = __builtin_amdgcn_raw_buffer_load_b64(rT, toff[c], so, 0);
                        d.t[ta][r][c] = {__uint_as_float(v.x), __uint_as_float(v.y)};
                    }
                }
            }
            // one state word per sample
            const v4u w01 = __builtin_amdgcn_raw_buffer_load_b128(rS, (16 * ta + 4 * g) * 8, 0, 0);
            const v4u w23 = __builtin_amdgcn_raw_buffer_load_b128(rS, (16 * ta + 4 * g + 2) * 8, 0, 0);
            const v2u wn = __builtin_amdgcn_raw_buffer_load_b64(rS, (16 * ta + n) * 8, 0, 0);
            d.ws[ta][0] = ((uint64_t)w01.y << 32) | w01.x;
            d.ws[ta][1] = ((uint64_t)w01.w << 32) | w01.z;
            d.ws[ta][2] = ((uint64_t)w23.y << 32) | w23.x;
            d.ws[ta][3] = ((uint64_t)w23.w << 32) | w23.z;
            d.wa[ta] = ((uint64_t)wn.y << 32) | wn.x;
        }
    };

    const int64_t nch = (a.S + kChunk - 1) / kChunk;
    const int64_t nwv = (int64_t)gridDim.x * NWB;
    // the chunk index is wave-uniform (scalar buffer resources)
    int64_t ch = (int64_t)blockIdx.x * NWB + __builtin_amdgcn_readfirstlane(w);
    auto process = [&](int64_t ch, const Chunk& cu) {
        const int64_t s0 = ch * kChunk;
        // sigma bits of samples 16 ta + 4 g + r at spins 16 rt + n
        unsigned sb[TA][4][RT];
#pragma unroll
        for (int ta = 0; ta < TA; ta++)
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int rt = 0; rt < RT; rt++)
                    sb[ta][r][rt] = (unsigned)(cu.ws[ta][r] >> (16 * rt + n)) & 1u;
        // ---- y of samples 16 ta + 4 g + r
        cplx<float> y[TA][4];
        if (mode == 0) {
            cplx<float> u[TA][4];
#pragma unroll
            for (int ta = 0; ta < TA; ta++)
#pragma unroll
                for (int r = 0; r < 4; r++) u[ta][r] = {0.f, 0.f};
            // stage-1 A operands: sample 16 ta + n, spins 32 kc + 8 g + j
            v4u A1[TA][KC];
#pragma unroll
            for (int ta = 0; ta < TA; ta++)
#pragma unroll
                for (int kc = 0; kc < KC; kc++) {
                    const unsigned e = (unsigned)(cu.wa[ta] >> (32 * kc + 8 * g)) & 0xffu;
#pragma unroll
                    for (int h = 0; h < 4; h++) A1[ta][kc][h] = pm1_pair((e >> (2 * h)) & 1u, (e >> (2 * h + 1)) & 1u);
                }
#pragma unroll
            for (int c = 0; c < CT; c++) {
                v4f gr[TA], gi[TA];
#pragma unroll
                for (int ta = 0; ta < TA; ta++) gr[ta] = gi[ta] = v4f{0, 0, 0, 0};
#pragma unroll
                for (int kc = 0; kc < KC; kc++) {
                    const v4u* xb = xs + (kc * CT + c) * 6 * 64 + lane;
                    const v4u rh = xb[0], rm = xb[64], rl = xb[128], ih = xb[192], im = xb[256], il = xb[320];
#pragma unroll
                    for (int ta = 0; ta < TA; ta++) {
                        gr[ta] = mfma_bf16(A1[ta][kc], rl, gr[ta]);
                        gi[ta] = mfma_bf16(A1[ta][kc], il, gi[ta]);
                        gr[ta] = mfma_bf16(A1[ta][kc], rm, gr[ta]);
                        gi[ta] = mfma_bf16(A1[ta][kc], im, gi[ta]);
                        gr[ta] = mfma_bf16(A1[ta][kc], rh, gr[ta]);
                        gi[ta] = mfma_bf16(A1[ta][kc], ih, gi[ta]);
                    }
                }
#pragma unroll
                for (int ta = 0; ta < TA; ta++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const cplx<float> t = cu.t[ta][r][c];
                        const float gre = gr[ta][r] + xh[c].re, gim = gi[ta][r] + xh[c].im;
                        u[ta][r].re += t.re * gre - t.im * gim;
                        u[ta][r].im += t.re * gim + t.im * gre;
                    }
            }
            // sigma . x_v over the lane's spins n + 16 rt
#pragma unroll
            for (int ta = 0; ta < TA; ta++)
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int rt = 0; rt < RT; rt++) {
                        const bool neg = sb[ta][r][rt] != 0;
                        u[ta][r].re += neg ? -xv[rt].re : xv[rt].re;
                        u[ta][r].im += neg ? -xv[rt].im : xv[rt].im;
                    }
#pragma unroll
            for (int ta = 0; ta < TA; ta++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const bool ok = s0 + 16 * ta + 4 * g + r < a.S;
                    const float ur = group_sum<16>(u[ta][r].re), ui = group_sum<16>(u[ta][r].im);
                    y[ta][r] = ok ? cplx<float>{ur, ui} : cplx<float>{0.f, 0.f};
                }
        } else {
#pragma unroll
            for (int ta = 0; ta < TA; ta++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t s = s0 + 16 * ta + 4 * g + r, sc = min(s, a.S - 1);
                    const cplx<float> v = mode == 1 ? cplx<float>{(float)a.obs[sc * 2 * a.n_groups],
                                                                  (float)a.obs[sc * 2 * a.n_groups + 1]}
                                                    : cplx<float>{1.f, 0.f};
                    y[ta][r] = s < a.S ? v : cplx<float>{0.f, 0.f};
                }
        }
        // ---- stage-2 A operands: spins 16 rt + n of the slot samples
        // (element j: sample 16 (j/4) + 4 g + j%4)
        using AV = typename StageTwo<TA>::V;
        AV A2[RT];
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int h = 0; h < 2 * TA; h++)
                A2[rt][h] = pm1_pair(sb[h / 2][2 * (h % 2)][rt], sb[h / 2][2 * (h % 2) + 1][rt]);
        // z_v += sigma^T y
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int ta = 0; ta < TA; ta++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const bool neg = sb[ta][r][rt] != 0;
                    zv[rt].re += neg ? -y[ta][r].re : y[ta][r].re;
                    zv[rt].im += neg ? -y[ta][r].im : y[ta][r].im;
                }
        // ---- stage 2 per column tile: b = y conj(t), split, Z += A2 b
#pragma unroll
        for (int c = 0; c < CT; c++) {
            float br[4 * TA], bi[4 * TA];
#pragma unroll
            for (int j = 0; j < 4 * TA; j++) {
                const cplx<float> t = cu.t[j / 4][j % 4][c], yy = y[j / 4][j % 4];
                br[j] = yy.re * t.re + yy.im * t.im;
                bi[j] = yy.im * t.re - yy.re * t.im;
                zh[c].re += br[j];
                zh[c].im += bi[j];
                if (mode == 2) t2[c] += t.re * t.re + t.im * t.im;  // t = 0 past S
            }
            AV Brh, Brm, Brl, Bih, Bim, Bil;
#pragma unroll
            for (int h = 0; h < 2 * TA; h++) {
                const Split3 sr = split3(br[2 * h], br[2 * h + 1]), si = split3(bi[2 * h], bi[2 * h + 1]);
                Brh[h] = sr.hi;
                Brm[h] = sr.mid;
                Brl[h] = sr.lo;
                Bih[h] = si.hi;
                Bim[h] = si.mid;
                Bil[h] = si.lo;
            }
#pragma unroll
            for (int rt = 0; rt < RT; rt++) {
                zr[rt][c] = StageTwo<TA>::mfma(A2[rt], Brl, zr[rt][c]);
                zi[rt][c] = StageTwo<TA>::mfma(A2[rt], Bil, zi[rt][c]);
                zr[rt][c] = StageTwo<TA>::mfma(A2[rt], Brm, zr[rt][c]);
                zi[rt][c] = StageTwo<TA>::mfma(A2[rt], Bim, zi[rt][c]);
                zr[rt][c] = StageTwo<TA>::mfma(A2[rt], Brh, zr[rt][c]);
                zi[rt][c] = StageTwo<TA>::mfma(A2[rt], Bih, zi[rt][c]);
            }
        }
    };
    // two chunk buffers alternate, one loaded while the other is used, where
    // the registers allow it (no spills up to CT = 4, RT * CT = 8).  The next
    // chunk is always requested (past the end: empty range, no memory
    // access): with a condition, the wait before using the current chunk
    // would have to cover the path without requests and so wait for all.
    constexpr bool PF = TA == 2 ? RT * CT <= 2 : RT * CT <= 4;
    Chunk b0;
    if constexpr (PF) {
        Chunk b1;
        fetch(ch, b0);
        while (ch < nch) {
            fetch(ch + nwv, b1);
            process(ch, b0);
            ch += nwv;
            if (ch >= nch) break;
            fetch(ch + nwv, b0);
            process(ch, b1);
            ch += nwv;
        }
    } else {
        for (; ch < nch; ch += nwv) {
            fetch(ch, b0);
            process(ch, b0);
        }
    }

    // ---- block reduction: a tree over the waves through LDS (fixed order),
    // then all threads write the row
    constexpr int NQ = C::NQ, RS = C::RS;
    float v[NQ];
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int c = 0; c < CT; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                v[((rt * CT + c) * 2) * 4 + r] = zr[rt][c][r];
                v[((rt * CT + c) * 2 + 1) * 4 + r] = zi[rt][c][r];
            }
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        v[C::oZV + 2 * rt] = zv[rt].re;
        v[C::oZV + 2 * rt + 1] = zv[rt].im;
    }
#pragma unroll
    for (int c = 0; c < CT; c++) {
        v[C::oZH + 2 * c] = zh[c].re;
        v[C::oZH + 2 * c + 1] = zh[c].im;
        v[C::oT2 + c] = t2[c];
    }
    float* red = reinterpret_cast<float*>(xs);
#pragma unroll
    for (int half = NWB / 2; half >= 1; half /= 2) {
        __syncthreads();  // the region (and, first, the X terms) is free
        if (w >= half && w < 2 * half) {
            float* dst = red + (w - half) * NQ * RS + lane;
#pragma unroll
            for (int q = 0; q < NQ; q++) dst[q * RS] = v[q];
        }
        __syncthreads();
        if (w < half) {
            const float* src = red + w * NQ * RS + lane;
            float o[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) o[q] = src[q * RS];
#pragma unroll
            for (int q = 0; q < NQ; q++) v[q] += o[q];
        }
    }
    __syncthreads();
    if (w == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) red[q * RS + lane] = v[q];
    __syncthreads();
    // block total of partial q over the four lane groups of column n
    auto gsum = [&](int q, int nn) {
        const float* p = red + q * RS + nn;
        return (p[0] + p[16]) + (p[32] + p[48]);
    };
    cplx<float>* out = a.part + (int64_t)blockIdx.x * a.ld;
    for (int p = tid; p < N + M; p += NT) {
        cplx<float> val;
        if (p < N) {
            const int rt = p / 16, nn = p % 16;
            val = {gsum(C::oZV + 2 * rt, nn), gsum(C::oZV + 2 * rt + 1, nn)};
        } else {
            int c, nn;
            C::inv(p - N, c, nn);
            val = {gsum(C::oZH + 2 * c, nn), gsum(C::oZH + 2 * c + 1, nn)};
        }
        out[p] = val;
    }
    // Z entry (j, i) at N + M + j + i N; (i, j) advance by NT entries per step
    {
        const int di = NT / N, dj = NT % N;
        int i = tid / N, j = tid % N;
        for (int e = tid; e < N * M; e += NT) {
            int c, nn;
            C::inv(i, c, nn);
            const int rt = j / 16, gg = (j % 16) / 4, r = j % 4, ln = 16 * gg + nn;
            const int q = ((rt * CT + c) * 2) * 4 + r;
            out[N + M + e] = {red[q * RS + ln], red[(q + 4) * RS + ln]};
            i += di;
            j += dj;
            if (j >= N) {
                j -= N;
                i += 1;
            }
        }
    }
    if (mode == 2)
        for (int i = tid; i < M; i += NT) {
            int c, nn;
            C::inv(i, c, nn);
            out[a.P + i] = {gsum(C::oT2 + c, nn), 0.f};
        }
}

}  // namespace rbmgpu
