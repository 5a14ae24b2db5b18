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
// in rounds of one block per CU); for the MFMA forward, when only one
// block fits (the weights in LDS), that variant with 16 chains per block
// (the full MFMA tile, half the rounds); otherwise the unconstrained one.
template <int K>
static hipError_t mlp_one(const LearnArgs& a, hipStream_t st) {
    size_t lds = lds_bytes<float>(a.s.d) + mlp_lds_bytes(a.Hn, a.Np, a.s.d.N, a.kmax);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return hipErrorInvalidDevice;
    const bool wf = mlp_wave_fwd(a.Hn, a.s.d.N);
    int cpb = kMlpChains;
    if (wf)
        for (cpb = 1; cpb < kMlpChains && (int64_t)cpb * cus < a.s.n_chains;) cpb *= 2;
    const int grid = (a.s.n_chains + cpb - 1) / cpb;
    if (grid > cus) {
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute((const void*)k_sample_mlp<K, 4>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        int per_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sample_mlp<K, 4>,
                                                                    64 * cpb, lds);
        if (e != hipSuccess) return e;
        if (per_cu >= 2) return mlp_launch<K, 4>(a, grid, cpb, lds, st);
        if (!wf) {
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sample_mlp<K, 4>,
                                                             64 * kMlpChainsMax, lds);
            if (e != hipSuccess) return e;
            if (per_cu >= 1)
                return mlp_launch<K, 4>(a, (a.s.n_chains + kMlpChainsMax - 1) / kMlpChainsMax, kMlpChainsMax, lds,
                                        st);
        }
    }
    return mlp_launch<K, 1>(a, grid, cpb, lds, st);
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
    // one record per block of the least power of two (>= 64) threads that
    // covers the hidden units and the sites (the block reductions halve it)
    int blk = 64;
    while (blk < a.Hn || blk < a.N) blk *= 2;
    hipLaunchKernelGGL(k_train_fwd, dim3((unsigned)a.B), dim3(blk), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // per batch slice: gW2 = dZ^T A (N x Hn) with gb2 = column sums of dZ,
    // gW1 = dU^T S (Hn x N) with gb1 = column sums of dU, into the slice's
    // row of part (gW1, gb1, gW2, gb2 as in the gradient); then the sum
    // over slices in slice order
    const int64_t G = 2 * (int64_t)a.Hn * a.N + a.Hn + a.N;
    const int64_t oW1 = 0, ob1 = (int64_t)a.Hn * a.N, oW2 = ob1 + a.Hn, ob2 = oW2 + (int64_t)a.N * a.Hn;
    hipLaunchKernelGGL(k_gemm_tn_part, dim3((a.Hn + 63) / 64, (a.N + 63) / 64, a.ks), dim3(256), 0, st, a.dZ, a.N,
                       a.A, a.Hn, a.part + oW2, G, ob2 - oW2, a.N, a.Hn, a.B, a.kc);
    hipLaunchKernelGGL(k_gemm_tn_part, dim3((a.N + 63) / 64, (a.Hn + 63) / 64, a.ks), dim3(256), 0, st, a.dU, a.Hn,
                       a.S, a.N, a.part + oW1, G, ob1 - oW1, a.Hn, a.N, a.B, a.kc);
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, st, a.part, a.ks, G, G, a.gW1);
    return hipGetLastError();
}

}  // namespace rbmgpu
