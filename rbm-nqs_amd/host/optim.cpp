// optim.cpp -- see optim.hpp.
#include "optim.hpp"

#include <cmath>

namespace optim {

bool step_ok(const zd* dw, size_t n, double* max_abs) {
    double mx = 0;
    bool finite = true;
    for (size_t i = 0; i < n; i++) {
        const double a = std::abs(dw[i]);
        if (!std::isfinite(a)) finite = false;
        else if (a > mx) mx = a;
    }
    if (max_abs) *max_abs = finite ? mx : NAN;
    return finite && mx <= 30.;
}

}  // namespace optim
