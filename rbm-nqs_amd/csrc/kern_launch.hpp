// kern_launch.hpp -- launchers that dispatch the runtime launch shape to the
// compiled template instances; included once per precision.
#pragma once
#include <hip/hip_ext.h>

#include "kern_impl.hpp"

namespace rbmgpu {

template <typename R, int L, int K, bool PF>
static hipError_t sample_one(const SampArgs<R>& a, int block, hipStream_t st) {
    int64_t threads = (int64_t)a.n_chains * L;
    int grid = (int)((threads + block - 1) / block);
    size_t lds = lds_bytes<R>(a.d);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sample<R, L, K, PF>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sample<R, L, K, PF>), dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}
template <typename R, int L, int K, bool PF>
static hipError_t eval_one(const EvalArgs<R>& a, int block, hipStream_t st) {
    int64_t threads = a.n * L;
    int grid = (int)((threads + block - 1) / block);
    size_t lds = lds_bytes<R>(a.d);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_eval<R, L, K, PF>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_eval<R, L, K, PF>), dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}

#define RBM_SHAPES(X) \
    X(8, 1) X(8, 2) X(8, 4) X(8, 8) X(8, 16) X(16, 2) X(16, 4) X(16, 8) X(16, 16) X(32, 2) \
    X(32, 4) X(32, 8) X(32, 9) X(32, 16) X(64, 1) X(64, 2) X(64, 4) X(64, 5) X(64, 8)
// shapes with a Pfaffian instance: one chain per wavefront
#define PF_SHAPES(X) X(64, 1) X(64, 2) X(64, 4) X(64, 8)

template <typename R>
hipError_t launch_sample(const SampArgs<R>& a, Shape s, int block, hipStream_t st) {
    if (a.d.pf_rows) {
        if (block != 64 * a.d.waves) return hipErrorInvalidValue;
#define X(L_, K_) \
    if (s.L == L_ && s.K == K_) return sample_one<R, L_, K_, true>(a, block, st);
        PF_SHAPES(X)
#undef X
        return hipErrorInvalidValue;
    }
#define X(L_, K_) \
    if (s.L == L_ && s.K == K_) return sample_one<R, L_, K_, false>(a, block, st);
    RBM_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}
template <typename R>
hipError_t launch_eval(const EvalArgs<R>& a, Shape s, int block, hipStream_t st) {
    if (a.d.pf_rows) {
        if (block != 64 * a.d.waves) return hipErrorInvalidValue;
#define X(L_, K_) \
    if (s.L == L_ && s.K == K_) return eval_one<R, L_, K_, true>(a, block, st);
        PF_SHAPES(X)
#undef X
        return hipErrorInvalidValue;
    }
#define X(L_, K_) \
    if (s.L == L_ && s.K == K_) return eval_one<R, L_, K_, false>(a, block, st);
    RBM_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_build_O(const OArgs<R>& a, hipStream_t st) {
    if (a.S < 1) return hipSuccess;
    if (a.d.N > 128 || a.d.nw > 2) return hipErrorInvalidValue;
    const size_t lds = build_O_lds<R>(a.d);
    if (lds > 65536) {
        hipError_t e =
            hipFuncSetAttribute((const void*)k_build_O<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int64_t nbatch = (a.S + kOBatch - 1) / kOBatch;
    const unsigned blocks = (unsigned)(nbatch < 2048 ? nbatch : 2048);
    hipLaunchKernelGGL((k_build_O<R>), dim3(blocks), dim3(256), lds, st, a);
    return hipGetLastError();
}
template <typename R>
hipError_t launch_o_partial(const MomArgs<R>& a, int blocks, hipStream_t st) {
    if (blocks < 1 || a.P < 1 || a.chunk < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_o_partial<R>), dim3(blocks, (unsigned)((a.P + 63) / 64)), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename R, int ZM, int WAVES = 4>
static hipError_t sx_one(const SxArgs<R>& a, int blocks, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    size_t lds = 2 * (size_t)a.ld * sizeof(cplx<R>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sx_partial<R, ZM, WAVES>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (e0 || e1)
        hipExtLaunchKernelGGL((k_sx_partial<R, ZM, WAVES>), dim3(blocks), dim3(64 * WAVES), (uint32_t)lds, st, e0,
                              e1, 0u, a);
    else
        hipLaunchKernelGGL((k_sx_partial<R, ZM, WAVES>), dim3(blocks), dim3(64 * WAVES), lds, st, a);
    return hipGetLastError();
}
template <typename R>
hipError_t launch_sx(const SxArgs<R>& a, int blocks, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    if (a.ld <= 128 * 2) return sx_one<R, 2, 16>(a, blocks, st, e0, e1);
    if (a.ld <= 128 * 4) return sx_one<R, 4, 16>(a, blocks, st, e0, e1);
    if (a.ld <= 128 * 8) return sx_one<R, 8>(a, blocks, st, e0, e1);
    if (a.ld <= 128 * 12) return sx_one<R, 12>(a, blocks, st, e0, e1);
    if (a.ld <= 128 * 18) return sx_one<R, 18>(a, blocks, st, e0, e1);
    if (a.ld <= 128 * 24) return sx_one<R, 24>(a, blocks, st, e0, e1);
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_sx_reduce(const SxArgs<R>& a, hipStream_t st) {
    hipLaunchKernelGGL((k_sx_reduce<R>), dim3((a.P + 255) / 256, a.nchunks), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_sx_colsum(const SxArgs<R>& a, hipStream_t st) {
    hipLaunchKernelGGL((k_sx_colsum<R>), dim3((a.P + kColBlock - 1) / kColBlock), dim3(64 * kColWaves), 0, st, a);
    return hipGetLastError();
}

#define RBM_INSTANTIATE(R)                                                                  \
    template hipError_t launch_sx_colsum<R>(const SxArgs<R>&, hipStream_t);               \
    template hipError_t launch_sx_reduce<R>(const SxArgs<R>&, hipStream_t);               \
    template hipError_t launch_sample<R>(const SampArgs<R>&, Shape, int, hipStream_t);    \
    template hipError_t launch_eval<R>(const EvalArgs<R>&, Shape, int, hipStream_t);      \
    template hipError_t launch_build_O<R>(const OArgs<R>&, hipStream_t);                  \
    template hipError_t launch_o_partial<R>(const MomArgs<R>&, int, hipStream_t);         \
    template hipError_t launch_sx<R>(const SxArgs<R>&, int, hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace rbmgpu
