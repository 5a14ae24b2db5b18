// optim.hpp -- host-side optimizer arithmetic on P-length parameter
// vectors: the massive-step guard of abstract_optimizer::optimize
// (abstract_optimizer.cpp:157), the update plugins (plugin.cpp:30-77) and
// the plain gradient of gradient_descent (gradient_descent.cpp:49-60).
// Plain C++ on the host: these are O(P) per epoch.
#pragma once
#include <complex>
#include <cstddef>
#include <vector>

namespace optim {

using zd = std::complex<double>;

// abstract_optimizer.cpp:157 rejects an update whose largest |dw| exceeds 30.
// A non-finite entry is rejected too: the reference's max over a vector
// with NaN is unspecified, and a NaN step would be written into the weights
// and the next checkpoint.
bool step_ok(const zd* dw, size_t n, double* max_abs = nullptr);

// base_plugin::apply (plugin.hpp): turns the solver's step into the update
// (learning rate included)
class plugin {
   public:
    virtual ~plugin() = default;
    virtual void apply(zd* dw, size_t n, double lr) = 0;
};

// adam_plugin (plugin.cpp:30-57): bias-corrected first moment and separate
// second moments of the real and imaginary parts
class adam : public plugin {
   public:
    adam(size_t n, double beta1, double beta2, double eps);
    void apply(zd* dw, size_t n, double lr) override;

   private:
    double b1_, b2_, eps_;
    size_t t_ = 1;
    std::vector<zd> m_;
    std::vector<double> wr_, wi_;
};

// momentum_plugin (plugin.cpp:59-77): m = a m + (1 - a) dw with a ramped
// up by `dialup` from 1e-3 (or constant when dialup = 1); the first call
// starts from m = dw
class momentum : public plugin {
   public:
    momentum(size_t n, double alpha, double dialup);
    void apply(zd* dw, size_t n, double lr) override;

   private:
    double alpha_, dialup_, dup_;
    std::vector<zd> m_;
};

// gradient_descent::gradient (gradient_descent.cpp:49-60): dw = <E O*> -
// <E> <O>*, real part divided by real_factor
void sgd_gradient(const zd* eo, const zd* o_mean, zd e_mean, double real_factor, zd* dw, size_t n);

}  // namespace optim
