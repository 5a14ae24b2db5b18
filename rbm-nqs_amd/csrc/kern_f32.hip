// FP32 instances of the VMC kernels (FP32 arithmetic, FP64 observables).
#include "kern_launch.hpp"
namespace rbmgpu {
RBM_INSTANTIATE(float)
}
