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

adam::adam(size_t n, double beta1, double beta2, double eps)
    : b1_(beta1), b2_(beta2), eps_(eps), m_(n, 0.), wr_(n, 0.), wi_(n, 0.) {}

void adam::apply(zd* dw, size_t n, double lr) {
    const double c1 = 1 - std::pow(b1_, (double)t_), c2 = 1 - std::pow(b2_, (double)t_);
    t_++;
    for (size_t i = 0; i < n; i++) {
        m_[i] = b1_ * m_[i] + (1 - b1_) * dw[i];
        wr_[i] = b2_ * wr_[i] + (1 - b2_) * (dw[i].real() * dw[i].real());
        wi_[i] = b2_ * wi_[i] + (1 - b2_) * (dw[i].imag() * dw[i].imag());
        const zd mh = m_[i] / c1;
        const double wr = wr_[i] / c2, wi = wi_[i] / c2;
        dw[i] = lr * zd(mh.real() / (std::sqrt(wr) + eps_), mh.imag() / (std::sqrt(wi) + eps_));
    }
}

momentum::momentum(size_t n, double alpha, double dialup)
    : alpha_(alpha), dialup_(dialup), dup_(dialup == 1. ? 1. : 1e-3), m_(n, 0.) {}

void momentum::apply(zd* dw, size_t n, double lr) {
    double ax = alpha_;
    bool zero = true;  // Eigen's isZero: every |m_i| <= 1e-12
    for (size_t i = 0; i < n && zero; i++) zero = std::abs(m_[i]) <= 1e-12;
    if (zero)
        for (size_t i = 0; i < n; i++) m_[i] = dw[i];
    if (dup_ < 1.) {
        dup_ *= dialup_;
        ax *= dup_;
    }
    for (size_t i = 0; i < n; i++) {
        m_[i] = ax * m_[i] + (1 - ax) * dw[i];
        dw[i] = lr * m_[i];
    }
}

void sgd_gradient(const zd* eo, const zd* o_mean, zd e_mean, double real_factor, zd* dw, size_t n) {
    for (size_t i = 0; i < n; i++) {
        const zd g = eo[i] - std::conj(o_mean[i]) * e_mean;
        dw[i] = zd(g.real() / real_factor, g.imag());
    }
}

}  // namespace optim
