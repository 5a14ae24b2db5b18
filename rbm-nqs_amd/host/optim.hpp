// optim.hpp -- host-side optimizer arithmetic on P-length parameter
// vectors: the massive-step guard of abstract_optimizer::optimize
// (abstract_optimizer.cpp:157) and the step transforms applied after the
// solve.  Plain C++ on the host: these are O(P) per epoch.
#pragma once
#include <complex>
#include <cstddef>

namespace optim {

using zd = std::complex<double>;

// abstract_optimizer.cpp:157 rejects an update whose largest |dw| exceeds 30.
// A non-finite entry is rejected too: the reference's max over a vector
// with NaN is unspecified, and a NaN step would be written into the weights
// and the next checkpoint.
bool step_ok(const zd* dw, size_t n, double* max_abs = nullptr);

}  // namespace optim
