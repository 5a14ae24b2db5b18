// Instances of the structured (basic RBM) S.x / moments kernel (bf16 MFMA at
// FP32 accuracy).
#include <hip/hip_ext.h>

#include "kern_struct.hpp"

namespace rbmgpu {

// (CT, RT): M <= 16 CT hidden units, N <= 16 RT spins.  A basic RBM has
// M = alpha N >= N, so only RT <= CT occurs; larger shapes materialise O.
#define STRUCT_SHAPES(X) X(1, 1) X(2, 1) X(2, 2) X(4, 1) X(4, 2) X(4, 4) X(8, 1) X(8, 2)

static bool pick(int N, int M, int& ct, int& rt) {
    ct = rt = 0;
    if (N < 1 || M < 1) return false;
    int best = 1 << 30;
#define X(C_, R_)                                                              \
    if (16 * C_ >= M && 16 * R_ >= N && C_ * R_ * 256 < best) {              \
        best = C_ * R_ * 256;                                                  \
        ct = C_;                                                               \
        rt = R_;                                                               \
    }
    STRUCT_SHAPES(X)
#undef X
    // column tiles are loaded in pairs of adjacent units (16 bytes): M even
    if (ct % 2 == 0 && M % 2 != 0) ct = rt = 0;
    return ct != 0;
}

static size_t lds_for(int ct, int rt) {
#define X(C_, R_) \
    if (ct == C_ && rt == R_) return StructCfg<C_, R_>::bytes;
    STRUCT_SHAPES(X)
#undef X
    return 0;
}

size_t struct_lds_bytes(int N, int M) {
    int ct, rt;
    if (!pick(N, M, ct, rt)) return 0;
    return lds_for(ct, rt);
}

template <int CT, int RT>
static int occupancy() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_struct<CT, RT>, StructCfg<CT, RT>::THREADS,
                                                     lds_for(CT, RT)) != hipSuccess) {
        (void)hipGetLastError();  // clear it: the launch that follows reports its own status
        return 1;
    }
    return nb > 0 ? nb : 1;
}

int struct_blocks_per_cu(int N, int M, int* waves_per_block) {
    int ct, rt;
    if (!pick(N, M, ct, rt)) return 0;
#define X(C_, R_)                                                                               \
    if (ct == C_ && rt == R_) {                                                                 \
        *waves_per_block = StructCfg<C_, R_>::NWB;                                           \
        return occupancy<C_, R_>();                                                          \
    }
    STRUCT_SHAPES(X)
#undef X
    return 0;
}

template <int CT, int RT>
static hipError_t one(const StructArgs& a, int blocks, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    size_t lds = lds_for(CT, RT);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_struct<CT, RT>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    // with events: timestamps of this dispatch itself (start and end of the kernel)
    if (e0 || e1)
        hipExtLaunchKernelGGL((k_struct<CT, RT>), dim3(blocks), dim3(StructCfg<CT, RT>::THREADS), (uint32_t)lds,
                              st, e0, e1, 0u, a);
    else
        hipLaunchKernelGGL((k_struct<CT, RT>), dim3(blocks), dim3(StructCfg<CT, RT>::THREADS), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_struct(const StructArgs& a, int blocks, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    int ct, rt;
    if (!pick(a.N, a.M, ct, rt)) return hipErrorInvalidValue;
#define X(C_, R_) \
    if (ct == C_ && rt == R_) return one<C_, R_>(a, blocks, st, e0, e1);
    STRUCT_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace rbmgpu
