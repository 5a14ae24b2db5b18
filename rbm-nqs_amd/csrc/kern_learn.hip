// Learned-proposal sampler and proposal-network training (FP32 sampler,
// bf16 MFMA forward); see kern_learn.hpp.
#include "kern_learn.hpp"

namespace rbmgpu {

template <int K, int WPE>
static hipError_t mlp_launch(const LearnArgs& a, int grid, int cpb, size_t lds, hipStream_t st) {
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sample_mlp<K, WPE>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sample_mlp<K, WPE>), dim3(grid), dim3(64 * cpb), lds, st, a);
    return hipGetLastError();
}

// Chains per block: 8 in lockstep for the MFMA forward; with the
// per-wavefront forward the fewest (a power of two) that still give at
// most one block per CU, so that few chains spread over all CUs instead of
// sharing SIMDs.  More blocks than CUs: the variant limited to 128 VGPRs
// when two of its blocks fit a CU (all chains resident at once instead of
// in rounds of one block per CU); otherwise the unconstrained one.
template <int K>
static hipError_t mlp_one(const LearnArgs& a, hipStream_t st) {
    size_t lds = lds_bytes<float>(a.s.d) + mlp_lds_bytes(a.Hn, a.Np, a.s.d.N, a.kmax);
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return hipErrorInvalidDevice;
    int cpb = kMlpChains;
    if (mlp_wave_fwd(a.Hn, a.s.d.N))
        for (cpb = 1; cpb < kMlpChains && (int64_t)cpb * cus < a.s.n_chains;) cpb *= 2;
    const int grid = (a.s.n_chains + cpb - 1) / cpb;
    if (grid > cus) {
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute((const void*)k_sample_mlp<K, 4>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sample_mlp<K, 4>,
                                                                    64 * cpb, lds);
        if (e != hipSuccess) return e;
    }
    return per_cu >= 2 ? mlp_launch<K, 4>(a, grid, cpb, lds, st) : mlp_launch<K, 1>(a, grid, cpb, lds, st);
}

hipError_t launch_sample_mlp(const LearnArgs& a, int K, hipStream_t st) {
    switch (K) {
        case 1: return mlp_one<1>(a, st);
        case 2: return mlp_one<2>(a, st);
        case 4: return mlp_one<4>(a, st);
        case 5: return mlp_one<5>(a, st);
        case 8: return mlp_one<8>(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mlp_logits(const LearnArgs& a, const uint64_t* states, int64_t n, float* z, hipStream_t st) {
    size_t lds = mlp_lds_bytes(a.Hn, a.Np, a.s.d.N, a.kmax);
    if (lds > 65536) {
        hipError_t e =
            hipFuncSetAttribute((const void*)k_mlp_logits, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int grid = (int)((n + kMlpChains - 1) / kMlpChains);
    hipLaunchKernelGGL(k_mlp_logits, dim3(grid), dim3(64 * kMlpChains), lds, st, a, states, n, z);
    return hipGetLastError();
}

hipError_t launch_train_grad(const TrainArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_train_fwd, dim3((unsigned)a.B), dim3(512), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // gW2 = dZ^T A (N x Hn), gW1 = dU^T S (Hn x N), biases = column sums
    hipLaunchKernelGGL(k_gemm_tn, dim3((a.Hn + 63) / 64, (a.N + 63) / 64), dim3(256), 0, st, a.dZ, a.N, a.A, a.Hn,
                       a.gW2, a.N, a.Hn, a.B);
    hipLaunchKernelGGL(k_gemm_tn, dim3((a.N + 63) / 64, (a.Hn + 63) / 64), dim3(256), 0, st, a.dU, a.Hn, a.S, a.N,
                       a.gW1, a.Hn, a.N, a.B);
    hipLaunchKernelGGL(k_colsum_f, dim3((a.Hn + 255) / 256), dim3(256), 0, st, a.dU, a.Hn, a.Hn, a.B, a.gb1);
    hipLaunchKernelGGL(k_colsum_f, dim3((a.N + 255) / 256), dim3(256), 0, st, a.dZ, a.N, a.N, a.B, a.gb2);
    return hipGetLastError();
}

}  // namespace rbmgpu
