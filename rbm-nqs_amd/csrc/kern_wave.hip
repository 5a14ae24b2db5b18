// FP32 one-chain-per-wavefront sampler instances (kern_wave.hpp).
#include "kern_wave.hpp"

namespace rbmgpu {

template <int K, bool SYM, bool PAIRS>
static hipError_t wave_one(const SampArgs<float>& a, int block, hipStream_t st) {
    const int waves = block / 64;
    const int grid = (a.n_chains + waves - 1) / waves;
    const size_t lds = wave_lds_bytes(a.d, K, a.npair);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sample_wave<K, SYM, PAIRS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sample_wave<K, SYM, PAIRS>), dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_sample_wave(const SampArgs<float>& a, int K, int block, hipStream_t st) {
    // K units per lane with 64 (K - 1) < H <= 64 K: only a lane's last unit
    // can be a padding unit
    if (block < 64 || block > 1024 || block % 64 || a.d.pf_rows || a.d.nw > 2 || a.d.H < 1 ||
        K != (a.d.H + 63) / 64)
        return hipErrorInvalidValue;
    const bool sym = a.d.n_symm > 1;
    if (sym ? 64 % a.d.n_alpha != 0 : a.npair != 0) return hipErrorInvalidValue;
#define WAVE_K(K_)                                                               \
    case K_:                                                                     \
        return !sym           ? wave_one<K_, false, false>(a, block, st)      \
               : a.npair > 0 ? wave_one<K_, true, true>(a, block, st)        \
                             : wave_one<K_, true, false>(a, block, st);
    switch (K) {
        WAVE_K(1)
        WAVE_K(2)
        WAVE_K(3)
        WAVE_K(4)
        WAVE_K(5)
        WAVE_K(6)
        WAVE_K(7)
        WAVE_K(8)
    }
#undef WAVE_K
    return hipErrorInvalidValue;
}

}  // namespace rbmgpu
