// adapters.hpp -- the reference's machine / sampler / operators / optimizer
// interfaces implemented on top of the GPU C ABI (include/rbmgpu.h), so the
// `rbm --train/--evaluate` driver keeps its structure (src/main.cpp).
//
//   machine::gpu_rbm            abstract_machine / rbm_base / rbm_symmetry
//   sampler::gpu_metropolis     abstract_sampler / metropolis_sampler
//   operators::aggregator       aggregator (finalised results of one sample())
//   optimizer::gpu_cg_solver    abstract_solver / cg_solver
//   optimizer::stochastic_reconfiguration   SR optimiser + decay_t
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/rbmgpu.h"
#include "setup.hpp"

namespace machine {

using cplx = std::complex<double>;

// spin configuration, bit i set <=> sigma_i = -1 (spin_state.hpp:34-47)
struct spin_state {
    size_t n;
    uint64_t w[2] = {0, 0};
    explicit spin_state(size_t n_) : n(n_) {}
    bool operator[](size_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
    void flip(size_t i) { w[i >> 6] ^= (uint64_t)1 << (i & 63); }
    void flip(const std::vector<size_t>& f) {
        for (size_t i : f) flip(i);
    }
};

class gpu_rbm {
   public:
    explicit gpu_rbm(const rbmhost::setup& s);
    ~gpu_rbm();
    gpu_rbm(const gpu_rbm&) = delete;
    gpu_rbm& operator=(const gpu_rbm&) = delete;

    const size_t n_visible;
    size_t get_n_params() const { return (size_t)n_params_; }
    size_t get_n_neural_params() const { return (size_t)n_neural_; }
    size_t get_n_updates() const { return n_updates_; }
    bool has_pfaffian() const { return n_params_ > n_neural_; }

    // rbm_base.cpp:58-113 (+ pfaffian.cpp:139-155 when present)
    void initialize_weights(std::mt19937& rng, double std_dev, double std_dev_imag = -1.,
                            double pf_std = 0.01);
    // w -= dw and count the update (rbm_base.cpp:117-143)
    void update_weights(const std::vector<cplx>& dw);
    void update_weights_nc(const std::vector<cplx>& dw);
    cplx log_psi(const spin_state& s);
    cplx psi(const spin_state& s) { return std::exp(log_psi(s)); }
    cplx psi_over_psi(const spin_state& s, const std::vector<size_t>& flips);
    // .rbm checkpoints: W, h_bias, v_bias as eigen_fstream matrices, then the
    // update counter, then the Pfaffian fs_ (rbm_base.cpp:219-288)
    bool save(const std::string& name, bool silent = false);
    bool load(const std::string& name);
    const std::vector<cplx>& params() const { return params_; }
    void set_params(const std::vector<cplx>& p);
    rbmgpu_ctx* ctx() { return ctx_; }

   private:
    const rbmhost::setup& setup_;
    rbmgpu_ctx* ctx_ = nullptr;
    int64_t n_params_ = 0, n_neural_ = 0;
    size_t n_updates_ = 0;
    std::vector<cplx> params_;
    void push();
};

}  // namespace machine

namespace operators {

// finalised statistics of one observable (aggregator.hpp:35-131)
struct aggregator {
    std::vector<std::complex<double>> result;
    std::vector<double> variance;
    double stddev = 0, tau = 0, variance_binned = 0;
    const std::vector<std::complex<double>>& get_result() const { return result; }
    const std::vector<double>& get_variance() const { return variance; }
    double get_stddev() const { return stddev; }
    double get_tau() const { return tau; }
};

}  // namespace operators

namespace sampler {

class gpu_metropolis_sampler {
   public:
    // evaluation: observables of --evaluate (hexagon plaquettes) and
    // sampler.eval_samples (main.cpp:640-680)
    gpu_metropolis_sampler(machine::gpu_rbm& rbm, const rbmhost::setup& s);
    // one epoch of all chains plus the finalisation of the registered
    // aggregators (metropolis_sampler.cpp:55-116, aggregator.cpp:84-164)
    void sample();
    double get_acceptance_rate() const { return m_.acceptance; }
    size_t get_n_samples() const { return n_samples_; }
    size_t get_my_n_samples() const { return n_samples_; }
    const operators::aggregator& energy() const { return a_h_; }
    const operators::aggregator& derivative() const { return a_d_; }
    const operators::aggregator& energy_derivative() const { return a_dh_; }
    const std::vector<std::complex<double>>& group_means() const { return groups_; }
    rbmgpu_ctx* ctx() { return rbm_.ctx(); }

   private:
    machine::gpu_rbm& rbm_;
    size_t n_samples_;
    uint32_t epoch_ = 0;
    rbmgpu_moments m_{};
    operators::aggregator a_h_, a_d_, a_dh_;
    std::vector<std::complex<double>> groups_;
};

}  // namespace sampler

namespace optimizer {

// decay_t of abstract_optimizer.cpp:27-54: get() decays before returning
struct decay_t {
    double initial, min, decay, value;
    bool min_reached;
    decay_t(const ini::decay_t& d, size_t offset);
    double get();
};

class gpu_cg_solver {
   public:
    gpu_cg_solver(rbmgpu_ctx* ctx, size_t n, size_t max_iterations, double rtol)
        : ctx_(ctx), n_(n), max_it_(max_iterations), rtol_(rtol) {}
    // abstract_solver::solve with O, <O>, norm and Var(O) resident on the GPU
    void solve(const std::vector<std::complex<double>>& b, std::vector<std::complex<double>>& x, double r1,
               double r2, double rd);
    int64_t get_itn() const { return itn_; }
    double get_rs() const { return rs_; }

   private:
    rbmgpu_ctx* ctx_;
    size_t n_, max_it_;
    double rtol_;
    int64_t itn_ = 0;
    double rs_ = 0;
};

class stochastic_reconfiguration {
   public:
    stochastic_reconfiguration(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s, const ini::config& cfg);
    // F = <O* E> - <E><O>* and the CG solve (SR.cpp:81-116)
    std::vector<std::complex<double>>& gradient(bool log);
    // abstract_optimizer.cpp:80-178 without the resample heuristic
    bool optimize();
    double get_current_energy() const { return sampler_.energy().result.at(0).real(); }
    const gpu_cg_solver& solver() const { return solver_; }
    double last_energy_per_site() const;
    std::vector<std::pair<std::string, double>> log_values;

   private:
    machine::gpu_rbm& rbm_;
    sampler::gpu_metropolis_sampler& sampler_;
    decay_t lr_, kp1_, kp2_, kp1d_;
    gpu_cg_solver solver_;
    std::vector<std::complex<double>> dw_, F_;
};

}  // namespace optimizer
