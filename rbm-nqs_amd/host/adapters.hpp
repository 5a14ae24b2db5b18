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

    // rbm_base.cpp:58-113 (+ pfaffian.cpp:139-155 when present), with the
    // configuration's rbm.weights / weights_imag / weights_type
    void initialize_weights(std::mt19937& rng);
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
    // the host copy, refreshed from the device after device-side updates
    const std::vector<cplx>& params();
    void set_params(const std::vector<cplx>& p);
    // a device-resident epoch updated the weights (rbmgpu_train_epoch)
    void note_device_update() {
        device_newer_ = true;
        n_updates_++;
    }
    rbmgpu_ctx* ctx() { return ctx_; }

   private:
    const rbmhost::setup& setup_;
    rbmgpu_ctx* ctx_ = nullptr;
    int64_t n_params_ = 0, n_neural_ = 0;
    size_t n_updates_ = 0;
    std::vector<cplx> params_;
    bool device_newer_ = false;
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
    // aggregators (metropolis_sampler.cpp:55-116, aggregator.cpp:84-164);
    // with a learned proposal, one training step of its network on the
    // epoch's records follows
    void sample();
    // the epoch index of the next sampling epoch (Philox stream), consumed
    uint32_t take_epoch() { return epoch_++; }
    // the energy statistics of a device-resident epoch (rbmgpu_train_epoch)
    void set_epoch_stats(const rbmgpu_epoch_stats& st);
    void train_proposal();
    double get_acceptance_rate() const { return m_.acceptance; }
    // sampler.type = full: every state of the basis, weighted by |psi|^2
    bool is_full() const { return full_; }
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
    bool full_ = false;
    int full_bits_ = 2;
    uint32_t epoch_ = 0;
    rbmgpu_moments m_{};
    operators::aggregator a_h_, a_d_, a_dh_;
    std::vector<std::complex<double>> groups_;
    int learn_hidden_ = 0;
    double learn_lr_ = 0, learn_entropy_ = 0;
    int64_t learn_batch_ = 0;
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

// The SR solver family of the reference on the device: "cg" and "minres"
// (the reference runs both as its distributed CG, cg_solver.cpp:31-37),
// "cg-single" (explicit S, cg_solver.cpp:274-299), "cg-dynamic" (the flop
// model of cg_solver.cpp:38-46 picks one of the two) and "direct"
// (direct_solver.cpp:31-59).  abstract_solver::solve with O, <O>, the
// sample count and Var(O) resident on the GPU.
class gpu_solver {
   public:
    gpu_solver(rbmgpu_ctx* ctx, size_t n, size_t n_samples, size_t max_iterations, double rtol,
               const std::string& method);
    void solve(const std::vector<std::complex<double>>& b, std::vector<std::complex<double>>& x, double r1,
               double r2, double rd);
    int32_t method() const { return method_; }
    size_t max_iterations() const { return max_it_; }
    double rtol() const { return rtol_; }
    int64_t get_itn() const { return itn_; }
    double get_rs() const { return rs_; }

   private:
    rbmgpu_ctx* ctx_;
    size_t n_, max_it_;
    double rtol_;
    int32_t method_;
    int64_t itn_ = 0;
    double rs_ = 0;
};

// base_plugin (plugin.hpp): the step transform applied instead of dw *= lr
class base_plugin {
   public:
    virtual ~base_plugin() = default;
    virtual void apply(std::vector<std::complex<double>>& dw, double lr) = 0;
};

class abstract_optimizer;

// heun_plugin (plugin.cpp:79-113): a half step, a new sample and gradient
// at the shifted weights, and a step size from the difference of the two
// gradients measured in the metric of the current samples
class heun_plugin : public base_plugin {
   public:
    heun_plugin(abstract_optimizer& opt, machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s, double eps)
        : opt_(opt), rbm_(rbm), sampler_(s), eps_(eps) {}
    void apply(std::vector<std::complex<double>>& dw, double lr) override;

   private:
    abstract_optimizer& opt_;
    machine::gpu_rbm& rbm_;
    sampler::gpu_metropolis_sampler& sampler_;
    double eps_;
};

// abstract_optimizer (abstract_optimizer.cpp:27-192): the learning-rate
// decay, the resample heuristic, the massive-step guard, the plugin and
// the weight update around a gradient of the subclass
class abstract_optimizer {
   public:
    abstract_optimizer(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s, const ini::config& cfg);
    virtual ~abstract_optimizer() = default;
    virtual std::vector<std::complex<double>>& gradient(bool log) = 0;
    // abstract_optimizer.cpp:80-178, after sampler.sample()
    bool optimize();
    // one whole epoch on the device (sampling included, rbmgpu_train_epoch)
    // where the configuration allows it; false if the guard refused the step
    virtual bool fused() const { return false; }
    virtual bool train_step() { return false; }
    void set_plugin(std::unique_ptr<base_plugin> p) { plug_ = std::move(p); }
    bool has_metric() const { return metric_; }
    double get_current_energy() const { return sampler_.energy().result.at(0).real(); }
    double last_energy_per_site() const { return get_current_energy() / (double)rbm_.n_visible; }
    size_t get_total_resamples() const { return total_resamples_; }
    std::vector<std::pair<std::string, double>> log_values;

   protected:
    machine::gpu_rbm& rbm_;
    sampler::gpu_metropolis_sampler& sampler_;
    decay_t lr_;
    bool resample_;
    double alpha1_, alpha2_, alpha3_;
    std::complex<double> last_energy_ = 0;
    double last_energy_std_ = -1;
    size_t total_resamples_ = 0;
    std::vector<std::complex<double>> dw_, last_update_;
    std::unique_ptr<base_plugin> plug_;
    bool metric_ = false;  // SR adds the sample metric for heun (SR.cpp:110-112)
    void log_energy(const char* std_key);
};

// stochastic_reconfiguration (SR.cpp:39-116)
class stochastic_reconfiguration : public abstract_optimizer {
   public:
    stochastic_reconfiguration(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s, const ini::config& cfg);
    std::vector<std::complex<double>>& gradient(bool log) override;
    bool fused() const override;
    bool train_step() override;
    const gpu_solver& solver() const { return solver_; }

   private:
    decay_t kp1_, kp2_, kp1d_;
    gpu_solver solver_;
    std::vector<std::complex<double>> F_;
    bool learned_ = false;
};

// gradient_descent (gradient_descent.cpp:24-60)
class gradient_descent : public abstract_optimizer {
   public:
    gradient_descent(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s, const ini::config& cfg);
    std::vector<std::complex<double>>& gradient(bool log) override;

   private:
    double real_factor_;
};

// init_optimizer (main.cpp:241-287): optimizer type and plugin from the
// configuration; throws for options the GPU backend does not run
std::unique_ptr<abstract_optimizer> make_optimizer(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s,
                                                   const ini::config& cfg);

}  // namespace optimizer
