// kern_struct.hpp -- S.x and O statistics of the basic RBM without the O
// matrix, as FP32 MFMA GEMMs (gfx950 v_mfma_f32_16x16x4_f32).
//
// For the basic RBM (rbm_base.cpp:194-217) the O column of sample k is
//   O_k = [ sigma_k (N) | t_k (M) | vec(sigma_k t_k^T) (N x M, j + i N) ],
// t_k = tanh theta_k, so with x = [x_v | x_h | X (N x M)]:
//   y_k = O_k . x = sigma_k . x_v + sum_i t_ki (x_hi + (Sigma X)_ki)
//   sum_k conj(O_k) y_k = [ Sigma^T y | conj(T)^T y | Sigma^T diag(y) conj(T) ].
// Sigma (S x N, entries +-1) and X / T are the only operands: one CG matvec
// reads sigma bits and T (S x M complex) once instead of O (S x P).  Per tile
// of 64 samples, stage 1 forms G = Sigma_tile X (real x complex = two real
// GEMMs), y by a 16-lane reduction of t (x_h + G) in the accumulator layout;
// stage 2 accumulates Z += Sigma_tile^T (diag(y) conj(T_tile)) in registers
// across all tiles of the block.  Modes: 0 = S.x (y from x), 1 = y = E_loc
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

// CT = Mp/16 column tiles (stage 1, per wave), RT = Np/16 row tiles (stage 2)
template <int CT, int RT>
__global__ void __launch_bounds__(256) k_struct(StructArgs a) {
    constexpr int Mp = 16 * CT, Np = 16 * RT;
    constexpr int PAIRS = RT * CT;            // (row tile, col tile) pairs of Z
    constexpr int PPW = (PAIRS + 3) / 4;      // pairs per wave
    extern __shared__ __align__(16) char smem[];
    float* Xr = reinterpret_cast<float*>(smem);  // Np x Mp (row = spin)
    float* Xi = Xr + Np * Mp;
    float* sg = Xi + Np * Mp;                    // kTS x Np (row = sample)
    cplx<float>* Tt = reinterpret_cast<cplx<float>*>(sg + kTS * Np);  // kTS x Mp
    cplx<float>* yv = Tt + kTS * Mp;             // kTS
    float* xv = reinterpret_cast<float*>(yv + kTS);  // 2 Np (re, im)
    float* xh = xv + 2 * Np;                         // 2 Mp
    if (a.status && a.status->done) return;
    const int N = a.N, M = a.M, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    if (a.mode == 0) {
        for (int e = tid; e < Np * Mp; e += blockDim.x) {
            int j = e / Mp, i = e % Mp;
            bool in = j < N && i < M;
            int64_t q = in ? (int64_t)N + M + j + (int64_t)i * N : 0;
            Xr[e] = in ? (float)a.x[2 * q] : 0.f;
            Xi[e] = in ? (float)a.x[2 * q + 1] : 0.f;
        }
        for (int j = tid; j < Np; j += blockDim.x) {
            xv[2 * j] = j < N ? (float)a.x[2 * j] : 0.f;
            xv[2 * j + 1] = j < N ? (float)a.x[2 * j + 1] : 0.f;
        }
        for (int i = tid; i < Mp; i += blockDim.x) {
            xh[2 * i] = i < M ? (float)a.x[2 * (N + i)] : 0.f;
            xh[2 * i + 1] = i < M ? (float)a.x[2 * (N + i) + 1] : 0.f;
        }
    }
    // stage-2 accumulators: wave w owns pairs w, w+4, ... (re and im)
    v4f zr[PPW], zi[PPW];
#pragma unroll
    for (int q = 0; q < PPW; q++) zr[q] = zi[q] = v4f{0, 0, 0, 0};
    // column sums kept by threads: z_v (spin j = tid < Np), z_h / |t|^2 (hidden i = tid < Mp)
    cplx<float> zv{0, 0}, zh{0, 0};
    float t2 = 0;
    const int64_t ntiles = (a.S + kTS - 1) / kTS;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t k0 = tile * kTS;
        __syncthreads();  // previous tile's LDS reads are done
        for (int e = tid; e < kTS * Np; e += blockDim.x) {
            int k = e / Np, j = e % Np;
            int64_t s = k0 + k;
            float v = 0.f;
            if (s < a.S && j < N) v = ((a.states[s * a.nw + (j >> 6)] >> (j & 63)) & 1) ? -1.f : 1.f;
            sg[e] = v;
        }
        for (int e = tid; e < kTS * Mp; e += blockDim.x) {
            int k = e / Mp, i = e % Mp;
            int64_t s = k0 + k;
            Tt[e] = (s < a.S && i < M) ? a.T[s * M + i] : cplx<float>{0.f, 0.f};
        }
        __syncthreads();
        // ---- y for the 64 samples of the tile
        if (a.mode == 0) {
            // stage 1: G = Sigma_tile X, rows 16w..16w+15 of the tile
            v4f gr[CT], gi[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) gr[c] = gi[c] = v4f{0, 0, 0, 0};
            for (int s4 = 0; s4 < Np / 4; s4++) {
                float av = sg[(16 * w + l16) * Np + 4 * s4 + lq];
#pragma unroll
                for (int c = 0; c < CT; c++) {
                    int bi = (4 * s4 + lq) * Mp + 16 * c + l16;
                    gr[c] = mfma16x16x4(av, Xr[bi], gr[c]);
                    gi[c] = mfma16x16x4(av, Xi[bi], gi[c]);
                }
            }
            // u = sum_i t_ki (x_hi + G_ki) for the 4 samples of this lane
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
            for (int r = 0; r < 4; r++) {
                ur[r] = group_sum<16>(ur[r]);
                ui[r] = group_sum<16>(ui[r]);
            }
            if (l16 < 4) {  // lane l16 == r writes sample 16w + 4 lq + r
                int r = l16;
                int k = 16 * w + 4 * lq + r;
                float vr = 0, vi = 0;  // sigma_k . x_v
                for (int j = 0; j < Np; j++) {
                    vr += sg[k * Np + j] * xv[2 * j];
                    vi += sg[k * Np + j] * xv[2 * j + 1];
                }
                float sr = r == 0 ? ur[0] : r == 1 ? ur[1] : r == 2 ? ur[2] : ur[3];
                float si = r == 0 ? ui[0] : r == 1 ? ui[1] : r == 2 ? ui[2] : ui[3];
                bool valid = k0 + k < a.S;
                yv[k] = valid ? cplx<float>{sr + vr, si + vi} : cplx<float>{0.f, 0.f};
            }
        } else {
            for (int k = tid; k < kTS; k += blockDim.x) {
                int64_t s = k0 + k;
                cplx<float> y{0.f, 0.f};
                if (s < a.S) y = a.mode == 1 ? cplx<float>{(float)a.obs[s * 2 * a.n_groups], (float)a.obs[s * 2 * a.n_groups + 1]}
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
            for (int s4 = 0; s4 < kTS / 4; s4++) {
                int k = 4 * s4 + lq;
                float av = sg[k * Np + 16 * rt + l16];
                cplx<float> t = Tt[k * Mp + 16 * ct + l16];
                cplx<float> y = yv[k];
                float br = y.re * t.re + y.im * t.im;  // y conj(t)
                float bi = y.im * t.re - y.re * t.im;
                zr[q] = mfma16x16x4(av, br, zr[q]);
                zi[q] = mfma16x16x4(av, bi, zi[q]);
            }
        }
        // ---- column sums: z_v = Sigma^T y, z_h = conj(T)^T y, sum |t|^2
        if (tid < Np) {
            for (int k = 0; k < kTS; k++) {
                float s = sg[k * Np + tid];
                zv.re += s * yv[k].re;
                zv.im += s * yv[k].im;
            }
        }
        if (tid >= 128 && tid - 128 < Mp) {
            int i = tid - 128;
            for (int k = 0; k < kTS; k++) {
                cplx<float> t = Tt[k * Mp + i], y = yv[k];
                zh.re += y.re * t.re + y.im * t.im;
                zh.im += y.im * t.re - y.re * t.im;
                t2 += t.re * t.re + t.im * t.im;
            }
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
    if (tid < N) out[tid] = zv;
    if (tid >= 128 && tid - 128 < M) {
        out[N + (tid - 128)] = zh;
        if (a.mode == 2) out[a.P + (tid - 128)] = {t2, 0.f};
    }
}

}  // namespace rbmgpu
