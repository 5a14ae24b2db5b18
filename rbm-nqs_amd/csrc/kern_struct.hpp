// kern_struct.hpp -- S.x and O statistics of the basic RBM without the O
// matrix, as FP32 MFMA GEMMs (gfx950 v_mfma_f32_16x16x4_f32).
//
// For the basic RBM (rbm_base.cpp:194-217) the O column of sample k is
//   O_k = [ sigma_k (N) | t_k (M) | vec(sigma_k t_k^T) (N x M, j + i N) ],
// t_k = tanh theta_k, so with x = [x_v | x_h | X (N x M)]:
//   y_k = O_k . x = sigma_k . x_v + sum_i t_ki (x_hi + (Sigma X)_ki)
//   sum_k conj(O_k) y_k = [ Sigma^T y | conj(T)^T y | Sigma^T diag(y) conj(T) ].
// Sigma (S x N, entries +-1) and X / T are the only operands: one CG matvec
// reads the sigma bits and T (S x M complex) once instead of O (S x P).
//
// A block of 4 waves walks tiles of 64 samples.  Per tile: stage 1 forms
// G = Sigma_tile X (real x complex = two real GEMMs; wave w owns samples
// 16w..16w+15), y by a 16-lane reduction of t (x_h + G) + sigma . x_v in the
// accumulator layout; stage 2 accumulates Z += Sigma_tile^T diag(y)
// conj(T_tile) in registers across all tiles of the block.  The column sums
// Sigma^T y and conj(T)^T y are spread over all 256 threads (each thread
// owns one column and a stride of samples) and reduced once at the end.
// The next tile's T and state words are loaded into registers while the
// current tile is computed.  Modes: 0 = S.x (y from x), 1 = y = E_loc
// (gives sum E conj(O)), 2 = y = 1 (gives sum conj(O) and sum |t|^2).
//
// MFMA operand layout (16x16x4, one block): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; accumulator register r of lane l is D[4(l/16)+r][l%16].
#pragma once
#include "dev_common.hpp"
#include "kern_api.hpp"

namespace rbmgpu {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline v4f mfma16x16x4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int kTS = 64;  // samples per tile

// LDS layout of k_struct<CT, RT> (floats): X re/im (Np x Mp each), sigma
// tile (kTS rows of stride Np + 1: odd, so the column reads of stage 2 and
// the row reads of stage 1 spread over the banks), T tile (kTS x Mp
// complex), y (kTS complex), x_v (2 Np), x_h (2 Mp)
template <int CT, int RT>
struct StructLds {
    static constexpr int Mp = 16 * CT, Np = 16 * RT, SP = Np + 1;
    static constexpr int oXr = 0, oXi = Np * Mp, oSg = 2 * Np * Mp, oT = oSg + ((kTS * SP + 1) & ~1),
                         oY = oT + 2 * kTS * Mp, oXv = oY + 2 * kTS, oXh = oXv + 2 * Np, total = oXh + 2 * Mp;
    static constexpr size_t bytes = (size_t)total * 4;
};

template <int CT, int RT>
__global__ void __launch_bounds__(256) k_struct(StructArgs a) {
    using Lay = StructLds<CT, RT>;
    constexpr int Mp = Lay::Mp, Np = Lay::Np, SP = Lay::SP;
    constexpr int PAIRS = RT * CT;        // (row tile, col tile) pairs of Z
    constexpr int PPW = (PAIRS + 3) / 4;  // pairs per wave
    constexpr int TPT = kTS * Mp / 256;   // T tile complex entries per thread
    constexpr int SPT = Np / 4;           // sigma entries per thread (4 threads per sample)
    extern __shared__ __align__(16) float smem_f[];
    float* Xr = smem_f + Lay::oXr;
    float* Xi = smem_f + Lay::oXi;
    float* sg = smem_f + Lay::oSg;
    cplx<float>* Tt = reinterpret_cast<cplx<float>*>(smem_f + Lay::oT);
    cplx<float>* yv = reinterpret_cast<cplx<float>*>(smem_f + Lay::oY);
    float* xv = smem_f + Lay::oXv;
    float* xh = smem_f + Lay::oXh;
    if (a.status && a.status->done) return;
    const int N = a.N, M = a.M, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    if (a.mode == 0) {
        for (int e = tid; e < Np * Mp; e += 256) {
            int j = e / Mp, i = e % Mp;
            bool in = j < N && i < M;
            int64_t q = in ? (int64_t)N + M + j + (int64_t)i * N : 0;
            Xr[e] = in ? (float)a.x[2 * q] : 0.f;
            Xi[e] = in ? (float)a.x[2 * q + 1] : 0.f;
        }
        for (int j = tid; j < Np; j += 256) {
            xv[2 * j] = j < N ? (float)a.x[2 * j] : 0.f;
            xv[2 * j + 1] = j < N ? (float)a.x[2 * j + 1] : 0.f;
        }
        for (int i = tid; i < Mp; i += 256) {
            xh[2 * i] = i < M ? (float)a.x[2 * (N + i)] : 0.f;
            xh[2 * i + 1] = i < M ? (float)a.x[2 * (N + i) + 1] : 0.f;
        }
    }
    // stage-2 accumulators: wave w owns pairs w, w+4, ... (re and im)
    v4f zr[PPW], zi[PPW];
#pragma unroll
    for (int q = 0; q < PPW; q++) zr[q] = zi[q] = v4f{0, 0, 0, 0};
    // column sums: z_v of spin column cv over samples kv0 + s * (256 / Np),
    // z_h (and |t|^2) of hidden column ch over samples kh0 + s * (256 / Mp)
    const int cv = tid % Np, kv0 = tid / Np, ch = tid % Mp, kh0 = tid / Mp;
    cplx<float> zv{0, 0}, zh{0, 0};
    float t2 = 0;
    // register prefetch of one tile: T entries e = tid + 256 j, state words
    // of sample tid / 4
    cplx<float> tp[TPT];
    uint64_t sw0 = 0, sw1 = 0;
    const int sk = tid >> 2, sq = tid & 3;
    auto fetch = [&](int64_t tile) {
        const int64_t k0 = tile * kTS;
#pragma unroll
        for (int j = 0; j < TPT; j++) {
            int e = tid + 256 * j, k = e / Mp, i = e % Mp;
            int64_t s = k0 + k;
            tp[j] = (s < a.S && i < M) ? a.T[s * M + i] : cplx<float>{0.f, 0.f};
        }
        int64_t s = k0 + sk;
        sw0 = s < a.S ? a.states[s * a.nw] : 0;
        sw1 = (s < a.S && a.nw > 1) ? a.states[s * a.nw + 1] : 0;
    };
    const int64_t ntiles = (a.S + kTS - 1) / kTS;
    int64_t tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        const int64_t k0 = tile * kTS;
        const bool valid_s = k0 + sk < a.S;
        __syncthreads();  // previous tile's LDS reads are done
#pragma unroll
        for (int j = 0; j < TPT; j++) Tt[tid + 256 * j] = tp[j];
#pragma unroll
        for (int u = 0; u < SPT; u++) {
            int jj = sq * SPT + u;
            uint64_t wd = jj < 64 ? sw0 : sw1;
            float v = (valid_s && jj < N) ? (((wd >> (jj & 63)) & 1) ? -1.f : 1.f) : 0.f;
            sg[sk * SP + jj] = v;
        }
        __syncthreads();
        if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);
        // ---- y for the 64 samples of the tile
        if (a.mode == 0) {
            // stage 1: G = Sigma_tile X, rows 16w..16w+15 of the tile
            v4f gr[CT], gi[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) gr[c] = gi[c] = v4f{0, 0, 0, 0};
#pragma unroll
            for (int s4 = 0; s4 < Np / 4; s4++) {
                float av = sg[(16 * w + l16) * SP + 4 * s4 + lq];
#pragma unroll
                for (int c = 0; c < CT; c++) {
                    int bi = (4 * s4 + lq) * Mp + 16 * c + l16;
                    gr[c] = mfma16x16x4(av, Xr[bi], gr[c]);
                    gi[c] = mfma16x16x4(av, Xi[bi], gi[c]);
                }
            }
            // u = sum_i t_ki (x_hi + G_ki) + sigma_k . x_v for the 4 samples
            // k = 16w + 4lq + r of this lane (columns l16 + 16c, l16 + 16c')
            float ur[4] = {0, 0, 0, 0}, ui[4] = {0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < CT; c++) {
                int i = 16 * c + l16;
                float hr = xh[2 * i], hi = xh[2 * i + 1];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int k = 16 * w + 4 * lq + r;
                    cplx<float> t = Tt[k * Mp + i];
                    float gre = gr[c][r] + hr, gim = gi[c][r] + hi;
                    ur[r] += t.re * gre - t.im * gim;
                    ui[r] += t.re * gim + t.im * gre;
                }
            }
#pragma unroll
            for (int c = 0; c < RT; c++) {
                int j = 16 * c + l16;
                float vr = xv[2 * j], vi = xv[2 * j + 1];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    float s = sg[(16 * w + 4 * lq + r) * SP + j];
                    ur[r] += s * vr;
                    ui[r] += s * vi;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                ur[r] = group_sum<16>(ur[r]);
                ui[r] = group_sum<16>(ui[r]);
            }
            if (l16 < 4) {  // lane l16 == r writes sample 16w + 4 lq + r
                int r = l16;
                int k = 16 * w + 4 * lq + r;
                float sr = r == 0 ? ur[0] : r == 1 ? ur[1] : r == 2 ? ur[2] : ur[3];
                float si = r == 0 ? ui[0] : r == 1 ? ui[1] : r == 2 ? ui[2] : ui[3];
                yv[k] = k0 + k < a.S ? cplx<float>{sr, si} : cplx<float>{0.f, 0.f};
            }
        } else {
            for (int k = tid; k < kTS; k += 256) {
                int64_t s = k0 + k;
                cplx<float> y{0.f, 0.f};
                if (s < a.S)
                    y = a.mode == 1 ? cplx<float>{(float)a.obs[s * 2 * a.n_groups], (float)a.obs[s * 2 * a.n_groups + 1]}
                                    : cplx<float>{1.f, 0.f};
                yv[k] = y;
            }
        }
        __syncthreads();
        // ---- stage 2: Z += Sigma_tile^T (diag(y) conj(T_tile))
#pragma unroll
        for (int q = 0; q < PPW; q++) {
            int pr = w + 4 * q;
            if (pr >= PAIRS) break;
            int rt = pr / CT, ct = pr % CT;
#pragma unroll 4
            for (int s4 = 0; s4 < kTS / 4; s4++) {
                int k = 4 * s4 + lq;
                float av = sg[k * SP + 16 * rt + l16];
                cplx<float> t = Tt[k * Mp + 16 * ct + l16];
                cplx<float> y = yv[k];
                float br = y.re * t.re + y.im * t.im;  // y conj(t)
                float bi = y.im * t.re - y.re * t.im;
                zr[q] = mfma16x16x4(av, br, zr[q]);
                zi[q] = mfma16x16x4(av, bi, zi[q]);
            }
        }
        // ---- column sums
        for (int k = kv0; k < kTS; k += 256 / Np) {
            float s = sg[k * SP + cv];
            cplx<float> y = yv[k];
            zv.re += s * y.re;
            zv.im += s * y.im;
        }
        for (int k = kh0; k < kTS; k += 256 / Mp) {
            cplx<float> t = Tt[k * Mp + ch], y = yv[k];
            zh.re += y.re * t.re + y.im * t.im;
            zh.im += y.im * t.re - y.re * t.im;
            t2 += t.re * t.re + t.im * t.im;
        }
    }
    // ---- write the block partial row [z_v | z_h | Z | sum|t|^2]
    cplx<float>* out = a.part + (int64_t)blockIdx.x * a.ld;
#pragma unroll
    for (int q = 0; q < PPW; q++) {
        int pr = w + 4 * q;
        if (pr >= PAIRS) break;
        int rt = pr / CT, ct = pr % CT;
        int i = 16 * ct + l16;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int j = 16 * rt + 4 * lq + r;
            if (j < N && i < M) out[N + M + j + (int64_t)i * N] = {zr[q][r], zi[q][r]};
        }
    }
    // column partials of the 256 threads -> LDS (T tile area) -> one sum per column
    __syncthreads();
    cplx<float>* rv = Tt;        // 256 z_v partials
    cplx<float>* rh = Tt + 256;  // 256 z_h partials
    float* rt2 = reinterpret_cast<float*>(Tt + 512);
    rv[tid] = zv;
    rh[tid] = zh;
    rt2[tid] = t2;
    __syncthreads();
    if (tid < N) {
        cplx<float> s{0, 0};
        for (int q = tid; q < 256; q += Np) s = s + rv[q];
        out[tid] = s;
    }
    if (tid < M) {
        cplx<float> s{0, 0};
        float s2 = 0;
        for (int q = tid; q < 256; q += Mp) {
            s = s + rh[q];
            s2 += rt2[q];
        }
        out[N + tid] = s;
        if (a.mode == 2) out[a.P + tid] = {s2, 0.f};
    }
}

}  // namespace rbmgpu
