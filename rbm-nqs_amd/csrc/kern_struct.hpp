// kern_struct.hpp -- S.x and O statistics of the basic RBM without the O
// matrix, as bf16 MFMA GEMMs (gfx950) at FP32 accuracy.
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
// Work split: each wave walks chunks of 16 samples (one MFMA tile) on its
// own; no barrier inside the loop.
//  * the chunk's T rows and state words stream from HBM straight into the
//    wave's LDS buffer (buffer_load ... lds, two buffers: the next chunk
//    loads while this one is computed) and are copied into registers in
//    the layout below; shapes whose buffers do not fit in LDS load into
//    registers directly.
//  * stage 1 (mode 0): G = Sigma_chunk X + 1 x_h^T (v_mfma_f32_16x16x32_bf16,
//    K = 32 spins).  A = the chunk's sigma rows, built from the state bits;
//    B = the X terms, staged once in LDS; the accumulators start at x_h.
//    Accumulator register r of lane l holds G[4(l/16)+r][column l%16].
//  * the lane holds exactly the t of that layout, forms sum_i t (x_h + G) +
//    sigma . x_v over its columns, and a 16-lane reduction gives y for the
//    samples 4(l/16)+r.
//  * stage 2: Z += Sigma^T diag(y) conj(T) (v_mfma_f32_16x16x16_bf16, K = the
//    16 samples): lane l supplies K slots 4(l/16)+j, j < 4 -- the samples
//    whose t and y it holds -- so B = y conj(t) is formed in registers.
//  * column sums Sigma^T y, conj(T)^T y and sum |t|^2 accumulate in VALU
//    registers; the waves of a block (one block per CU) add their partials
//    in a fixed tree through LDS and write one partial row per block.
// Hidden-unit order: column n of column tile c is hidden unit
// 32(c/2) + 2n + c%2 (CT even), so one 16-byte access holds a lane's
// entries of tiles c = 2p and 2p + 1.
// Modes: 0 = S.x (y from x), 1 = y = E_loc (gives sum E conj(O)),
// 2 = y = 1 (gives sum conj(O) and sum |t|^2).
//
// MFMA operand layouts: lane l supplies A[l%16][K0(l/16)+j] and
// B[K0(l/16)+j][l%16] with K0 = 8 (16x16x32, j < 8) or 4 (16x16x16, j < 4);
// accumulator register r of lane l is D[4(l/16)+r][l%16].
#pragma once
#include "dev_common.hpp"
#include "kern_api.hpp"

namespace rbmgpu {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v4s __attribute__((ext_vector_type(4)));

// fourth word of a raw buffer resource on gfx9 (32-bit data format; the
// range check makes loads past num_records return 0)
constexpr int kBufWord3 = 0x00020000;
// s_waitcnt immediate (gfx9 fields): vmcnt 0, expcnt and lgkmcnt not waited for
constexpr int kWaitVm0 = 0x0F70;

// D += A B, 16x16x32, eight bf16 per lane and operand
__device__ inline v4f mfma_k32(v4u a, v4u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                   0);
}
// D += A B, 16x16x16, four bf16 per lane and operand
__device__ inline v4f mfma_k16(v2u a, v2u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(v4s, a), __builtin_bit_cast(v4s, b), c, 0, 0,
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

constexpr int kChunk = 16;  // samples per wave step

template <int CT, int RT>
struct StructCfg {
    static constexpr int Mp = 16 * CT, Np = 16 * RT;
    static constexpr int KC = (Np + 31) / 32;  // stage-1 K chunks of 32 spins
    static constexpr int CP = CT % 2 == 0 ? 2 : 1;
    // per-lane partial floats: Z (re, im), z_v, z_h, sum |t|^2
    static constexpr int oZV = RT * CT * 8, oZH = oZV + 2 * RT, oT2 = oZH + 2 * CT, NQ = oT2 + CT;
    static constexpr int NQ4 = (NQ + 3) / 4 * 4;  // padded to whole 16-byte LDS accesses
    // waves per SIMD the registers are budgeted for: two (256 registers)
    // hide each other's latencies; larger shapes (RT CT > 8, or CT = 8 with
    // its chunk in registers) would spill at two.  A block holds the waves
    // of one CU (4 SIMDs), so its partial row covers as many samples as
    // possible.
    static constexpr int WPE = RT * CT <= 8 && CT <= 4 ? 2 : 1;
    static constexpr int NWB = 4 * WPE, THREADS = 64 * NWB;
    static constexpr size_t xbytes = (size_t)KC * CT * 6 * 64 * 16;
    // direct global -> LDS chunk loads: per wave two buffers, each the
    // chunk's T rows (16 x Mp complex) and its 16 state words (in a 256-byte
    // slot), after the X terms; only where they fit in the 160 KB of LDS
    static constexpr size_t WT = (size_t)kChunk * Mp * 8, WB = WT + 256;
    static constexpr size_t dbytes = xbytes + (size_t)NWB * 2 * WB;
    static constexpr bool DMA = dbytes <= 163840;
    // tree reduction over the block's waves: NWB / 2 regions of NQ4 floats x 64 lanes
    static constexpr size_t rbytes = (size_t)(NWB / 2) * NQ4 * 64 * 4;
    static constexpr size_t lbytes = DMA ? dbytes : xbytes;
    static constexpr size_t bytes = lbytes > rbytes ? lbytes : rbytes;
    // hidden unit of column n of column tile c, and its inverse
    __host__ __device__ static constexpr int hid(int c, int n) { return 16 * CP * (c / CP) + CP * n + c % CP; }
    __device__ static void inv(int i, int& c, int& n) {
        const int p = i / (16 * CP), rem = i % (16 * CP);
        n = rem / CP;
        c = CP * p + rem % CP;
    }
};

template <int CT, int RT>
__global__ void __launch_bounds__((StructCfg<CT, RT>::THREADS), (StructCfg<CT, RT>::WPE)) k_struct(StructArgs a) {
    using C = StructCfg<CT, RT>;
    constexpr int NWB = C::NWB, NT = C::THREADS, KC = C::KC, CP = C::CP;
    static_assert(RT <= 4, "one state word per sample (N <= 64)");
    extern __shared__ __align__(16) v4u xs[];  // [KC][CT][6][64] X terms, chunk buffers, later the reduction
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

    // One chunk's operands in registers: t of samples 4 g + r (all columns
    // of the lane), their state words, and the word of sample n.  Loads carry
    // no condition (a condition puts each load behind a branch whose join
    // waits for every load in flight): rows past S lie outside the buffer
    // resource's range and read as 0, and those samples also get y = 0.  A
    // padded hidden unit (hid >= M) reads unit M - CP instead; it meets X = 0
    // and x_h = 0 in stage 1 and its Z, z_h and sum |t|^2 entries are not
    // written.
    struct Chunk {
        cplx<float> t[4][CT];
        uint64_t ws[4];
        uint64_t wa;
    };
    // byte offset of the lane's units in row 4 g (+ r) of a chunk
    uint32_t toff[CT / CP];
#pragma unroll
    for (int c = 0; c < CT; c += CP) toff[c / CP] = (uint32_t)(4 * g * M + min(C::hid(c, n), M - CP)) * 8u;
    // buffer resources of a chunk: its T rows and state words inside [0, S)
    // (empty for a chunk past the end: its loads return 0 without a memory
    // access)
    auto resources = [&](int64_t chn, __amdgpu_buffer_rsrc_t& rT, __amdgpu_buffer_rsrc_t& rS) {
        const int64_t s0 = chn * kChunk;
        const int rows = (int)max(min(a.S - s0, (int64_t)kChunk), (int64_t)0);
        rT = __builtin_amdgcn_make_buffer_rsrc(const_cast<cplx<float>*>(a.T + s0 * M), 0, rows * M * 8, kBufWord3);
        rS = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(a.states + s0), 0, rows * 8, kBufWord3);
    };

    const int64_t nch = (a.S + kChunk - 1) / kChunk;
    const int64_t nwv = (int64_t)gridDim.x * NWB;
    // the chunk index is wave-uniform (scalar buffer resources, LDS base in M0)
    const int wu = __builtin_amdgcn_readfirstlane(w);
    int64_t ch = (int64_t)blockIdx.x * NWB + wu;
    auto process = [&](int64_t chn, const Chunk& cu) {
        const int64_t s0 = chn * kChunk;
        // sigma bits of samples 4 g + r at spins 16 rt + n
        unsigned sb[4][RT];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int rt = 0; rt < RT; rt++) sb[r][rt] = (unsigned)(cu.ws[r] >> (16 * rt + n)) & 1u;
        // ---- y of samples 4 g + r
        cplx<float> y[4];
        if (mode == 0) {
            cplx<float> u[4];
#pragma unroll
            for (int r = 0; r < 4; r++) u[r] = {0.f, 0.f};
            // stage-1 A operands: sample n, spins 32 kc + 8 g + j
            v4u A1[KC];
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                const unsigned e = (unsigned)(cu.wa >> (32 * kc + 8 * g)) & 0xffu;
#pragma unroll
                for (int h = 0; h < 4; h++) A1[kc][h] = pm1_pair((e >> (2 * h)) & 1u, (e >> (2 * h + 1)) & 1u);
            }
#pragma unroll
            for (int c = 0; c < CT; c++) {
                v4f gr = {xh[c].re, xh[c].re, xh[c].re, xh[c].re};
                v4f gi = {xh[c].im, xh[c].im, xh[c].im, xh[c].im};
#pragma unroll
                for (int kc = 0; kc < KC; kc++) {
                    const v4u* xb = xs + (kc * CT + c) * 6 * 64 + lane;
                    const v4u rh = xb[0], rm = xb[64], rl = xb[128], ih = xb[192], im = xb[256], il = xb[320];
                    gr = mfma_k32(A1[kc], rl, gr);
                    gi = mfma_k32(A1[kc], il, gi);
                    gr = mfma_k32(A1[kc], rm, gr);
                    gi = mfma_k32(A1[kc], im, gi);
                    gr = mfma_k32(A1[kc], rh, gr);
                    gi = mfma_k32(A1[kc], ih, gi);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const cplx<float> t = cu.t[r][c];
                    u[r].re += t.re * gr[r] - t.im * gi[r];
                    u[r].im += t.re * gi[r] + t.im * gr[r];
                }
            }
            // sigma . x_v over the lane's spins n + 16 rt
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    const bool neg = sb[r][rt] != 0;
                    u[r].re += neg ? -xv[rt].re : xv[rt].re;
                    u[r].im += neg ? -xv[rt].im : xv[rt].im;
                }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool ok = s0 + 4 * g + r < a.S;
                const float ur = group_sum<16>(u[r].re), ui = group_sum<16>(u[r].im);
                y[r] = ok ? cplx<float>{ur, ui} : cplx<float>{0.f, 0.f};
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t s = s0 + 4 * g + r, sc = min(s, a.S - 1);
                const cplx<float> v = mode == 1 ? cplx<float>{(float)a.obs[sc * 2 * a.n_groups],
                                                              (float)a.obs[sc * 2 * a.n_groups + 1]}
                                                : cplx<float>{1.f, 0.f};
                y[r] = s < a.S ? v : cplx<float>{0.f, 0.f};
            }
        }
        // ---- stage-2 A operands: spins 16 rt + n of samples 4 g + j
        v2u A2[RT];
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            A2[rt][0] = pm1_pair(sb[0][rt], sb[1][rt]);
            A2[rt][1] = pm1_pair(sb[2][rt], sb[3][rt]);
        }
        // z_v += sigma^T y
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool neg = sb[r][rt] != 0;
                zv[rt].re += neg ? -y[r].re : y[r].re;
                zv[rt].im += neg ? -y[r].im : y[r].im;
            }
        // ---- stage 2 per column tile: b = y conj(t), split, Z += A2 b
#pragma unroll
        for (int c = 0; c < CT; c++) {
            float br[4], bi[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const cplx<float> t = cu.t[r][c], yy = y[r];
                br[r] = yy.re * t.re + yy.im * t.im;
                bi[r] = yy.im * t.re - yy.re * t.im;
                zh[c].re += br[r];
                zh[c].im += bi[r];
                if (mode == 2) t2[c] += t.re * t.re + t.im * t.im;  // t = 0 past S
            }
            const Split3 r01 = split3(br[0], br[1]), r23 = split3(br[2], br[3]);
            const Split3 i01 = split3(bi[0], bi[1]), i23 = split3(bi[2], bi[3]);
#pragma unroll
            for (int rt = 0; rt < RT; rt++) {
                zr[rt][c] = mfma_k16(A2[rt], v2u{r01.lo, r23.lo}, zr[rt][c]);
                zi[rt][c] = mfma_k16(A2[rt], v2u{i01.lo, i23.lo}, zi[rt][c]);
                zr[rt][c] = mfma_k16(A2[rt], v2u{r01.mid, r23.mid}, zr[rt][c]);
                zi[rt][c] = mfma_k16(A2[rt], v2u{i01.mid, i23.mid}, zi[rt][c]);
                zr[rt][c] = mfma_k16(A2[rt], v2u{r01.hi, r23.hi}, zr[rt][c]);
                zi[rt][c] = mfma_k16(A2[rt], v2u{i01.hi, i23.hi}, zi[rt][c]);
            }
        }
    };

    Chunk cu;
    if constexpr (C::DMA) {
        // The chunk streams from HBM into this wave's LDS buffer by
        // buffer_load ... lds (lane l's 16 bytes land at base + 16 l, so each
        // instruction copies 1 KB of consecutive T bytes; lanes past the
        // resource's range write zeros).  Per step: wait for the chunk in
        // flight (the only loads outstanding), copy it from LDS into
        // registers, start the next chunk into the other buffer, compute.
        typedef __attribute__((address_space(3))) void* lds_ptr;
        char* wbuf = reinterpret_cast<char*>(xs) + C::xbytes + (size_t)wu * 2 * C::WB;
        auto issue = [&](int64_t chn, char* dst) {
            __amdgpu_buffer_rsrc_t rT, rS;
            resources(chn, rT, rS);
#pragma unroll
            for (int j = 0; j < (int)(C::WT / 1024); j++)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rT, (lds_ptr)(dst + j * 1024), 16, lane * 16, j * 1024, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rS, (lds_ptr)(dst + C::WT), 4, lane * 4, 0, 0, 0);
        };
        auto copy_in = [&](const char* src) {
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < CT; c += CP) {
                    const char* at = src + r * M * 8 + toff[c / CP];
                    if constexpr (CP == 2) {
                        const float4 v = *reinterpret_cast<const float4*>(at);
                        cu.t[r][c] = {v.x, v.y};
                        cu.t[r][c + 1] = {v.z, v.w};
                    } else {
                        cu.t[r][c] = *reinterpret_cast<const cplx<float>*>(at);
                    }
                }
            const uint64_t* sw = reinterpret_cast<const uint64_t*>(src + C::WT);
#pragma unroll
            for (int r = 0; r < 4; r++) cu.ws[r] = sw[4 * g + r];
            cu.wa = sw[n];
        };
        if (ch < nch) issue(ch, wbuf);
        for (int k = 0; ch < nch; k ^= 1, ch += nwv) {
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            copy_in(wbuf + k * C::WB);
            if (ch + nwv < nch) issue(ch + nwv, wbuf + (k ^ 1) * C::WB);
            process(ch, cu);
        }
        __builtin_amdgcn_s_waitcnt(kWaitVm0);  // no load into LDS in flight: the reduction reuses it
    } else {
        // raw buffer loads into registers: the chunk base and the row
        // offsets are wave-uniform (scalar resource and offset), the lane
        // offsets are fixed
        for (; ch < nch; ch += nwv) {
            __amdgpu_buffer_rsrc_t rT, rS;
            resources(ch, rT, rS);
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < CT; c += CP) {
                    if constexpr (CP == 2) {
                        const v4u v = __builtin_amdgcn_raw_buffer_load_b128(rT, toff[c / 2], r * M * 8, 0);
                        cu.t[r][c] = {__uint_as_float(v.x), __uint_as_float(v.y)};
                        cu.t[r][c + 1] = {__uint_as_float(v.z), __uint_as_float(v.w)};
                    } else {
                        const v2u v = __builtin_amdgcn_raw_buffer_load_b64(rT, toff[c], r * M * 8, 0);
                        cu.t[r][c] = {__uint_as_float(v.x), __uint_as_float(v.y)};
                    }
                }
            const v4u w01 = __builtin_amdgcn_raw_buffer_load_b128(rS, 4 * g * 8, 0, 0);
            const v4u w23 = __builtin_amdgcn_raw_buffer_load_b128(rS, (4 * g + 2) * 8, 0, 0);
            const v2u wn = __builtin_amdgcn_raw_buffer_load_b64(rS, n * 8, 0, 0);
            cu.ws[0] = ((uint64_t)w01.y << 32) | w01.x;
            cu.ws[1] = ((uint64_t)w01.w << 32) | w01.z;
            cu.ws[2] = ((uint64_t)w23.y << 32) | w23.x;
            cu.ws[3] = ((uint64_t)w23.w << 32) | w23.z;
            cu.wa = ((uint64_t)wn.y << 32) | wn.x;
            process(ch, cu);
        }
    }

    // ---- block reduction: a tree over the waves through LDS (fixed order),
    // then all threads write the row.  Region layout: float q of lane l at
    // ((q / 4) 64 + l) 4 + q % 4 (16-byte accesses).
    constexpr int NQ4 = C::NQ4;
    float v[NQ4];
#pragma unroll
    for (int q = 0; q < NQ4; q++) v[q] = 0.f;
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
    float4* red4 = reinterpret_cast<float4*>(xs);
#pragma unroll
    for (int half = NWB / 2; half >= 1; half /= 2) {
        __syncthreads();  // the region (and, first, the X terms) is free
        if (w >= half && w < 2 * half) {
            float4* dst = red4 + (w - half) * (NQ4 / 4) * 64 + lane;
#pragma unroll
            for (int q = 0; q < NQ4; q += 4) dst[(q / 4) * 64] = float4{v[q], v[q + 1], v[q + 2], v[q + 3]};
        }
        __syncthreads();
        if (w < half) {
            const float4* src = red4 + w * (NQ4 / 4) * 64 + lane;
            float4 o[NQ4 / 4];
#pragma unroll
            for (int q = 0; q < NQ4; q += 4) o[q / 4] = src[(q / 4) * 64];
#pragma unroll
            for (int q = 0; q < NQ4; q += 4) {
                v[q] += o[q / 4].x;
                v[q + 1] += o[q / 4].y;
                v[q + 2] += o[q / 4].z;
                v[q + 3] += o[q / 4].w;
            }
        }
    }
    __syncthreads();
    if (w == 0)
#pragma unroll
        for (int q = 0; q < NQ4; q += 4) red4[(q / 4) * 64 + lane] = float4{v[q], v[q + 1], v[q + 2], v[q + 3]};
    __syncthreads();
    auto at = [&](int q, int ln) { return red[((q >> 2) * 64 + ln) * 4 + (q & 3)]; };
    // block total of partial q over the four lane groups of column n
    auto gsum = [&](int q, int nn) { return (at(q, nn) + at(q, nn + 16)) + (at(q, nn + 32) + at(q, nn + 48)); };
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
            out[N + M + e] = {at(q, ln), at(q + 4, ln)};
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
