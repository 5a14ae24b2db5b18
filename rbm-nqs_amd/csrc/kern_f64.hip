// FP64 instances of the VMC kernels (the parity mode).
#include "kern_launch.hpp"
namespace rbmgpu {
RBM_INSTANTIATE(double)
}
