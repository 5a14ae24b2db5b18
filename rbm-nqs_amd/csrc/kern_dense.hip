// kern_dense.hip -- the explicit S of the cg-single and direct SR solvers
// (cg_solver.cpp:274-299, direct_solver.cpp:31-59) and the dense S.x of
// the CG on it.
//
// S_ij = (1/S) sum_k conj(O_ki) O_kj - conj(mu_i) mu_j (+ regularisation).
// O is stored one sample per row (ld complex), i.e. as the real S x 2 ld
// matrix X with X[k][2p] = Re O_kp, X[k][2p+1] = Im O_kp, so G = X^T X is a
// real SYRK and
//   Re S_ij ~ G[2i][2j] + G[2i+1][2j+1],  Im S_ij ~ G[2i][2j+1] - G[2i+1][2j].
// k_syrk_*: one block per upper-triangle 64 x 64 tile of G and K slice
// (sample range).  Blocks are mapped XCD-aware: the dispatcher hands block
// b to XCD b % 8, and every tile of one K slice runs on the same XCD, so
// the slice's O rows come from HBM into that XCD's L2 once and the other
// tiles re-read them there.  FP32 on MFMA: 16-sample chunks (20 KB of LDS
// per block, 8 blocks per CU) of both panels
// go through registers into a double-buffered LDS stage (the global loads
// of chunk c + 1 are in flight while chunk c is multiplied); each of the 4
// waves owns a 32 x 32 quarter of the tile as 2 x 2
// v_mfma_f32_16x16x4_f32 tiles.  FP64 by VALU FMAs (4 x 4 per thread).
// Each K slice writes its own partial tile (no atomics, a fixed summation
// order); k_g_sum adds the slices in FP64 and k_s_from_g forms S.  The
// columns are centred by <O> while they are staged, so the products form
// the covariance directly: S = <O^H O> - <O>^H <O> would cancel
// catastrophically in FP32 where |<O>|^2 >> Var(O) (saturated units of
// converged states).
//
// MFMA 16x16x4 operand layout: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator r of lane l is D[4(l/16)+r][l%16].
#include <hip/hip_runtime.h>

#include "kern_api.hpp"

namespace rbmgpu {

typedef float v4f_d __attribute__((ext_vector_type(4)));

constexpr int kGT = 64;     // G tile edge
constexpr int kGChunk = 32; // samples per LDS chunk (FP64)
constexpr int kFChunk = 16; // samples per LDS chunk (FP32; 20 KB of LDS per block)
constexpr int kFRows = kFChunk / 16;  // chunk rows a thread stages
constexpr int kGRow = 80;   // FP32 LDS row stride: the 4 rows one MFMA operand read spans sit 16 banks apart

// tile pair q of the upper triangle (ti <= tj), row-major over ti
__device__ inline void pair_of(int q, int nt, int& ti, int& tj) {
    ti = 0;
    int rem = q;
    while (rem >= nt - ti) {
        rem -= nt - ti;
        ti++;
    }
    tj = ti + rem;
}

// block -> (tile pair q, K slice): with ksplit a multiple of 8 the K slices
// are dealt to the 8 XCDs round-robin and all npairs tiles of a slice share
// one XCD (block b runs on XCD b % 8)
__device__ inline void syrk_work(int npairs, int ksplit, int& q, int& slice) {
    const int b = blockIdx.x;
    if ((ksplit & 7) == 0) {
        const int x = b & 7, r = b >> 3;
        slice = (r / npairs) * 8 + x;
        q = r % npairs;
    } else {
        q = b % npairs;
        slice = b / npairs;
    }
}

// FP64: the 64 x 64 tile's panels for samples kc .. kc + 31 into LDS (zeros
// past the slice end and past the valid columns)
__device__ inline void syrk_stage_f64(const SyrkArgs<double>& a, double (*As)[kGT + 1], double (*Bs)[kGT + 1],
                                      int c0, int c1, int64_t kc, int64_t k0, int64_t k1) {
#pragma unroll
    for (int q = 0; q < kGChunk * kGT / 256; q++) {
        const int e = threadIdx.x + 256 * q, s = e / kGT, c = e % kGT;
        const int64_t k = kc + s;
        const bool in = k < k1;
        const int64_t row = (in ? k : k0) * a.ncol;
        const double f = a.wsx ? sqrt(a.wsx[in ? k : k0]) : 1.0;
        const bool va = in && c0 + c < a.nvalid, vb = in && c1 + c < a.nvalid;
        const double ma = a.mean && va ? a.mean[c0 + c] : 0.0, mb = a.mean && vb ? a.mean[c1 + c] : 0.0;
        As[s][c] = va ? f * (a.X[row + c0 + c] - ma) : 0.0;
        Bs[s][c] = vb ? f * (a.X[row + c1 + c] - mb) : 0.0;
    }
}

// FP32 staging: thread t carries columns 4 (t % 16) .. +3 of both panels for
// chunk rows t / 16 and t / 16 + 16
struct SyrkFetch {
    float4 a[kFRows], b[kFRows];
};
__device__ inline float4 centred4(const float* X, int64_t row, int col, int ncol, int nvalid, const float* m,
                                  float f) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (col < ncol) v = *reinterpret_cast<const float4*>(X + row + col);
    v.x = col + 0 < nvalid ? f * (v.x - m[0]) : 0.f;
    v.y = col + 1 < nvalid ? f * (v.y - m[1]) : 0.f;
    v.z = col + 2 < nvalid ? f * (v.z - m[2]) : 0.f;
    v.w = col + 3 < nvalid ? f * (v.w - m[3]) : 0.f;
    return v;
}
__device__ inline void syrk_fetch(const SyrkArgs<float>& a, SyrkFetch& r, int ca, int cb, bool diag,
                                  const float* ma, const float* mb, int64_t kc, int64_t k1) {
#pragma unroll
    for (int h = 0; h < kFRows; h++) {
        const int64_t k = kc + (threadIdx.x >> 4) + 16 * h;
        if (k < k1) {
            const int64_t row = k * a.ncol;
            const float f = a.wsx ? (float)sqrt(a.wsx[k]) : 1.f;
            r.a[h] = centred4(a.X, row, ca, a.ncol, a.nvalid, ma, f);
            r.b[h] = diag ? r.a[h] : centred4(a.X, row, cb, a.ncol, a.nvalid, mb, f);
        } else {
            r.a[h] = r.b[h] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

// FP32: MFMA, each wave a 32 x 32 quarter of the tile (2 x 2 MFMA tiles)
__global__ void __launch_bounds__(256) k_syrk_f32(SyrkArgs<float> a) {
    __shared__ __align__(16) float As[2][kFChunk][kGRow];
    __shared__ __align__(16) float Bs[2][kFChunk][kGRow];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
    const int wi = w >> 1, wj = w & 1;
    int q, slice, ti, tj;
    syrk_work(a.npairs, a.ksplit, q, slice);
    pair_of(q, a.ntile, ti, tj);
    const bool diag = ti == tj;
    const int64_t k0 = (int64_t)slice * a.kper, k1 = k0 + a.kper < a.S ? k0 + a.kper : a.S;
    const int cq = 4 * (tid & 15), ca = ti * kGT + cq, cb = tj * kGT + cq;
    float ma[4], mb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        ma[i] = a.mean && ca + i < a.nvalid ? (float)a.mean[ca + i] : 0.f;
        mb[i] = a.mean && cb + i < a.nvalid ? (float)a.mean[cb + i] : 0.f;
    }
    v4f_d acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = v4f_d{0, 0, 0, 0};
    SyrkFetch r;
    if (k0 < k1) {
        syrk_fetch(a, r, ca, cb, diag, ma, mb, k0, k1);
#pragma unroll
        for (int h = 0; h < kFRows; h++) {
            *reinterpret_cast<float4*>(&As[0][(tid >> 4) + 16 * h][cq]) = r.a[h];
            *reinterpret_cast<float4*>(&Bs[0][(tid >> 4) + 16 * h][cq]) = r.b[h];
        }
    }
    __syncthreads();
    int buf = 0;
    for (int64_t kc = k0; kc < k1; kc += kFChunk) {
        const bool more = kc + kFChunk < k1;
        if (more) syrk_fetch(a, r, ca, cb, diag, ma, mb, kc + kFChunk, k1);
#pragma unroll
        for (int kk = 0; kk < kFChunk; kk += 4) {
            const float a0 = As[buf][kk + lq][wi * 32 + lr], a1 = As[buf][kk + lq][wi * 32 + 16 + lr];
            const float b0 = Bs[buf][kk + lq][wj * 32 + lr], b1 = Bs[buf][kk + lq][wj * 32 + 16 + lr];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (more) {
            // the other buffer was last read before the previous barrier
#pragma unroll
            for (int h = 0; h < kFRows; h++) {
                *reinterpret_cast<float4*>(&As[buf ^ 1][(tid >> 4) + 16 * h][cq]) = r.a[h];
                *reinterpret_cast<float4*>(&Bs[buf ^ 1][(tid >> 4) + 16 * h][cq]) = r.b[h];
            }
        }
        __syncthreads();
        buf ^= 1;
    }
    float* out = a.part + ((int64_t)slice * a.npairs + q) * (kGT * kGT);
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int i = wi * 32 + x * 16 + 4 * lq + rr, j = wj * 32 + y * 16 + lr;
                out[i * kGT + j] = acc[x][y][rr];
            }
}

// FP64: VALU FMAs (MI355X runs FP64 vector and matrix FMAs at the same
// peak); thread t owns rows 4 (t / 16) .. +3 and columns t % 16 + 16 c
__global__ void __launch_bounds__(256) k_syrk_f64(SyrkArgs<double> a) {
    __shared__ double As[kGChunk][kGT + 1];
    __shared__ double Bs[kGChunk][kGT + 1];
    const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
    int q, slice, ti, tj;
    syrk_work(a.npairs, a.ksplit, q, slice);
    pair_of(q, a.ntile, ti, tj);
    const int64_t k0 = (int64_t)slice * a.kper, k1 = k0 + a.kper < a.S ? k0 + a.kper : a.S;
    double acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = 0;
    for (int64_t kc = k0; kc < k1; kc += kGChunk) {
        syrk_stage_f64(a, As, Bs, ti * kGT, tj * kGT, kc, k0, k1);
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kGChunk; k++) {
            double av[4], bv[4];
#pragma unroll
            for (int x = 0; x < 4; x++) av[x] = As[k][4 * ty + x];
#pragma unroll
            for (int y = 0; y < 4; y++) bv[y] = Bs[k][tx + 16 * y];
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) acc[x][y] = fma(av[x], bv[y], acc[x][y]);
        }
        __syncthreads();
    }
    double* out = a.part + ((int64_t)slice * a.npairs + q) * (kGT * kGT);
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) out[(4 * ty + x) * kGT + tx + 16 * y] = acc[x][y];
}

// G = the K slices' partial tiles summed in FP64, in slice order
template <typename R>
__global__ void __launch_bounds__(256) k_g_sum(SyrkArgs<R> a) {
    const int64_t n = (int64_t)a.npairs * kGT * kGT;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double s = 0;
        for (int k = 0; k < a.ksplit; k++) s += (double)a.part[(int64_t)k * n + e];
        a.G[e] = s;
    }
}

// G[u][v] (u, v < 2 ld; symmetric) from the summed tiles
template <typename R>
__device__ inline double g_at(const SyrkArgs<R>& a, int u, int v) {
    int tu = u / kGT, tv = v / kGT;
    if (tu > tv) {
        const int t = u;
        u = v;
        v = t;
        const int tt = tu;
        tu = tv;
        tv = tt;
    }
    const int q = tu * a.ntile - tu * (tu - 1) / 2 + (tv - tu);
    return a.G[(int64_t)q * (kGT * kGT) + (u % kGT) * kGT + (v % kGT)];
}

// S from G: thread (i, j) of the P x P complex matrix (column-major),
// mode 0: + r1 Var(O) (+ rd on the Pfaffian block) + r2 max Var on the
// diagonal (cg_solver.cpp:276-281); mode 1: the diagonal scaled by 1 + r1
// (1 + r1 + rd on the Pfaffian block), then + r2 max Var (direct_solver.cpp:45-53)
template <typename R>
__global__ void __launch_bounds__(256) k_s_from_g(SyrkArgs<R> a, int P, int n_neural, int64_t S, const double* mean,
                                                  const double* var, const double* maxvar, double r1, double r2,
                                                  double rd, const double* reg_dev, int mode, double* Sm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)P * P) return;
    if (reg_dev) {  // r1, r2, r1 delta from device memory (graph replays)
        r1 = reg_dev[0];
        r2 = reg_dev[1];
        rd = reg_dev[2];
    }
    const int i = (int)(t % P), j = (int)(t / P);
    const double inv = 1.0 / (double)S;
    double re = (g_at(a, 2 * i, 2 * j) + g_at(a, 2 * i + 1, 2 * j + 1)) * inv;
    double im = (g_at(a, 2 * i, 2 * j + 1) - g_at(a, 2 * i + 1, 2 * j)) * inv;
    if (!a.mean) {  // uncentred products: subtract conj(mu_i) mu_j
        const double mri = mean[2 * i], mii = mean[2 * i + 1], mrj = mean[2 * j], mij = mean[2 * j + 1];
        re -= mri * mrj + mii * mij;
        im -= mri * mij - mii * mrj;
    }
    if (i == j) {
        const double f = i < n_neural ? r1 : r1 + rd;
        if (mode == 0) re += f * var[i] + r2 * (*maxvar);
        else re = re * (1.0 + f) + r2 * (*maxvar);
        if (mode == 1) im *= 1.0 + f;
    }
    Sm[2 * t] = re;
    Sm[2 * t + 1] = im;
}

__global__ void __launch_bounds__(256) k_dense_mv(const double* Sm, const double* x, double* y, int P, double* pdot,
                                                  const CgStatus* st) {
    if (st && st->done) return;
    dense_mv_rows(Sm, x, y, P, pdot);
}

template <typename R>
static hipError_t g_sum(const SyrkArgs<R>& a, hipStream_t st) {
    const int64_t n = (int64_t)a.npairs * kGT * kGT;
    hipLaunchKernelGGL(k_g_sum<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}
template <>
hipError_t launch_syrk<float>(const SyrkArgs<float>& a, hipStream_t st) {
    if (a.ntile < 1 || a.ksplit < 1 || a.kper % kGChunk || a.ncol % 4 || !a.G) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_syrk_f32, dim3((unsigned)(a.npairs * a.ksplit)), dim3(256), 0, st, a);
    if (hipError_t e = hipGetLastError()) return e;
    return g_sum(a, st);
}
template <>
hipError_t launch_syrk<double>(const SyrkArgs<double>& a, hipStream_t st) {
    if (a.ntile < 1 || a.ksplit < 1 || a.kper % kGChunk || !a.G) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_syrk_f64, dim3((unsigned)(a.npairs * a.ksplit)), dim3(256), 0, st, a);
    if (hipError_t e = hipGetLastError()) return e;
    return g_sum(a, st);
}

template <typename R>
hipError_t launch_s_from_g(const SyrkArgs<R>& a, int P, int n_neural, int64_t S, const double* mean,
                           const double* var, const double* maxvar, double r1, double r2, double rd,
                           const double* reg_dev, int mode, double* Sm, hipStream_t st) {
    const int64_t n = (int64_t)P * P;
    hipLaunchKernelGGL(k_s_from_g<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, P, n_neural, S, mean,
                       var, maxvar, r1, r2, rd, reg_dev, mode, Sm);
    return hipGetLastError();
}

hipError_t launch_dense_mv(const double* Sm, const double* x, double* y, int P, double* pdot, const CgStatus* st,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_dense_mv, dim3((unsigned)((P + 3) / 4)), dim3(256), 0, s, Sm, x, y, P, pdot, st);
    return hipGetLastError();
}

#define DENSE_INST(R)                                                                                           \
    template hipError_t launch_s_from_g<R>(const SyrkArgs<R>&, int, int, int64_t, const double*, const double*, \
                                           const double*, double, double, double, const double*, int, double*,    \
                                           hipStream_t);
DENSE_INST(float)
DENSE_INST(double)

}  // namespace rbmgpu
