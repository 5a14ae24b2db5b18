// adapters.cpp -- see adapters.hpp.
#include "adapters.hpp"

#include "optim.hpp"

#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace {

void check(rbmgpu_ctx* c, int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + rbmgpu_last_error(c));
}

// eigen_fstream.cpp:25-51: [u64 rows][u64 cols][rows*cols complex128, column-major]
void write_mat(std::ofstream& f, const std::complex<double>* d, uint64_t rows, uint64_t cols) {
    f.write((const char*)&rows, 8);
    f.write((const char*)&cols, 8);
    f.write((const char*)d, rows * cols * sizeof(std::complex<double>));
}
bool read_mat(std::ifstream& f, std::complex<double>* d, uint64_t rows, uint64_t cols) {
    uint64_t r = 0, c = 0;
    f.read((char*)&r, 8);
    f.read((char*)&c, 8);
    if (!f || r != rows || c != cols) return false;
    f.read((char*)d, rows * cols * sizeof(std::complex<double>));
    return (bool)f;
}

}  // namespace

namespace machine {

gpu_rbm::gpu_rbm(const rbmhost::setup& s) : n_visible((size_t)s.layout().n_spins), setup_(s) {
    rbmgpu_config c = s.gpu_config();
    int rc = rbmgpu_create(&c, &ctx_);
    if (rc != 0) {
        std::string m = ctx_ ? rbmgpu_last_error(ctx_) : "rbmgpu_create failed";
        rbmgpu_destroy(ctx_);
        ctx_ = nullptr;
        throw std::runtime_error(m);
    }
    check(ctx_, rbmgpu_n_params(ctx_, &n_params_, &n_neural_), "n_params");
    params_.assign(n_params_, 0.);
}

gpu_rbm::~gpu_rbm() { rbmgpu_destroy(ctx_); }

void gpu_rbm::push() { check(ctx_, rbmgpu_set_params(ctx_, (const double*)params_.data()), "set_params"); }

void gpu_rbm::set_params(const std::vector<cplx>& p) {
    if ((int64_t)p.size() != n_params_) throw std::runtime_error("parameter vector size mismatch");
    params_ = p;
    device_newer_ = false;
    push();
}

void gpu_rbm::initialize_weights(std::mt19937& rng) {
    n_updates_ = 0;
    // rbm_base::initialize_weights (rbm_base.cpp:58-113), then the Pfaffian
    set_params(setup_.initial_weights(rng));
}

const std::vector<cplx>& gpu_rbm::params() {
    if (device_newer_) {
        check(ctx_, rbmgpu_get_params(ctx_, (double*)params_.data()), "get_params");
        device_newer_ = false;
    }
    return params_;
}

void gpu_rbm::update_weights(const std::vector<cplx>& dw) {
    params();
    for (int64_t i = 0; i < n_params_; i++) params_[i] -= dw[i];
    check(ctx_, rbmgpu_update_weights(ctx_, (const double*)dw.data()), "update_weights");
    n_updates_++;
}

void gpu_rbm::update_weights_nc(const std::vector<cplx>& dw) {
    update_weights(dw);
    n_updates_--;
}

cplx gpu_rbm::log_psi(const spin_state& s) {
    cplx out;
    check(ctx_, rbmgpu_eval_logpsi(ctx_, s.w, 1, (double*)&out), "eval_logpsi");
    return out;
}

cplx gpu_rbm::psi_over_psi(const spin_state& s, const std::vector<size_t>& flips) {
    std::vector<int32_t> f(flips.begin(), flips.end());
    cplx out;
    check(ctx_, rbmgpu_psi_ratio(ctx_, s.w, f.data(), (int32_t)f.size(), (double*)&out), "psi_ratio");
    return out;
}

bool gpu_rbm::save(const std::string& name, bool silent) {
    params();
    std::ofstream out{name + ".rbm", std::ios::binary};
    if (!out.is_open()) return false;
    const auto& L = setup_.layout();
    const int64_t nvb = L.n_vb, na = L.n_alpha, N = L.n_spins;
    write_mat(out, params_.data() + nvb + na, N, na);  // weights_ (N x n_alpha)
    write_mat(out, params_.data() + nvb, na, 1);       // h_bias_
    write_mat(out, params_.data(), nvb, 1);            // v_bias_
    uint64_t nu = n_updates_;
    out.write((const char*)&nu, 8);
    if (has_pfaffian()) write_mat(out, params_.data() + n_neural_, L.pf_rows, 4);
    if (!silent) std::cout << "Saved RBM to '" << name << ".rbm'!" << std::endl;
    return (bool)out;
}

bool gpu_rbm::load(const std::string& name) {
    std::ifstream in{name + ".rbm", std::ios::binary};
    if (!in.good()) return false;
    const auto& L = setup_.layout();
    const int64_t nvb = L.n_vb, na = L.n_alpha, N = L.n_spins;
    std::vector<cplx> p(n_params_, 0.);
    bool ok = read_mat(in, p.data() + nvb + na, N, na) && read_mat(in, p.data() + nvb, na, 1) &&
              read_mat(in, p.data(), nvb, 1);
    uint64_t nu = 0;
    in.read((char*)&nu, 8);
    ok = ok && (bool)in;
    if (ok && has_pfaffian()) ok = read_mat(in, p.data() + n_neural_, L.pf_rows, 4);
    if (!ok) throw std::runtime_error("Unable to load Eigen Matrix: dimension mismatch!");
    n_updates_ = nu;
    set_params(p);
    std::cout << "Loaded RBM from '" << name << ".rbm'!" << std::endl;
    return true;
}

}  // namespace machine

namespace sampler {

gpu_metropolis_sampler::gpu_metropolis_sampler(machine::gpu_rbm& rbm, const rbmhost::setup& s)
    : rbm_(rbm),
      n_samples_(s.cfg().sa_type == "full" ? (size_t)1 << rbm.n_visible : s.n_samples()),
      full_(s.cfg().sa_type == "full"),
      full_bits_((int)s.cfg().sa_full_n_parallel_bits),
      groups_(s.terms().n_groups),
      learn_hidden_(s.cfg().learn_hidden),
      learn_lr_(s.cfg().learn_lr),
      learn_entropy_(s.cfg().learn_entropy),
      learn_batch_(s.cfg().learn_batch) {}

void gpu_metropolis_sampler::train_proposal() {
    if (learn_hidden_ <= 0) return;
    double out[4];
    check(rbm_.ctx(), rbmgpu_learn_train(rbm_.ctx(), learn_lr_, learn_entropy_, learn_batch_, out), "learn_train");
}

void gpu_metropolis_sampler::set_epoch_stats(const rbmgpu_epoch_stats& st) {
    m_.e_mean[0] = st.e_mean[0];
    m_.e_mean[1] = st.e_mean[1];
    m_.e_var = st.e_var;
    m_.e_var_binned = st.e_var_binned;
    m_.e_tau = st.e_tau;
    m_.e_std = st.e_std;
    m_.n_bins = st.n_bins;
    m_.acceptance = st.acceptance;
    a_h_.result.assign(1, {st.e_mean[0], st.e_mean[1]});
    a_h_.variance.assign(1, st.e_var);
    a_h_.variance_binned = st.e_var_binned;
    a_h_.stddev = st.e_std;
    a_h_.tau = st.e_tau;
}

void gpu_metropolis_sampler::sample() {
    rbmgpu_ctx* c = rbm_.ctx();
    if (full_) check(c, rbmgpu_full_epoch(c, full_bits_), "full_epoch");  // full_sampler.cpp:43-185
    else check(c, rbmgpu_sample_epoch(c, epoch_++), "sample_epoch");
    int64_t P = (int64_t)rbm_.get_n_params();
    a_d_.result.resize(P);
    a_d_.variance.resize(P);
    a_dh_.result.resize(P);
    a_h_.result.resize(1);
    a_h_.variance.resize(1);
    check(c,
          rbmgpu_reduce(c, &m_, (double*)a_d_.result.data(), a_d_.variance.data(), (double*)a_dh_.result.data(),
                        (double*)groups_.data()),
          "reduce");
    a_h_.result[0] = {m_.e_mean[0], m_.e_mean[1]};
    a_h_.variance[0] = m_.e_var;
    a_h_.variance_binned = m_.e_var_binned;
    a_h_.stddev = m_.e_std;
    a_h_.tau = m_.e_tau;
    train_proposal();
}

}  // namespace sampler

namespace optimizer {

decay_t::decay_t(const ini::decay_t& d, size_t offset)
    : initial(d.initial), min(d.min), decay(d.decay), value(std::pow(d.decay, (double)offset) * d.initial),
      min_reached(d.decay == 1.) {}

double decay_t::get() {
    if (!min_reached) {
        value *= decay;
        if (value < min) {
            min_reached = true;
            value = min;
        }
    }
    return value;
}

namespace {
// cg_solver.cpp:31-46: "cg" and "minres" run the distributed CG;
// "cg-dynamic" compares the flop counts 2 n^2 (m + t) - n (n + t) (explicit
// S) and 4 n m t - t (m + n) (matrix-free) for n parameters, m samples and
// t = min(2n, max_iterations) iterations
int32_t solver_method(const std::string& s, size_t n, size_t m, size_t t) {
    if (s == "cg" || s == "minres") return RBMGPU_SOLVE_CG;
    if (s == "cg-single") return RBMGPU_SOLVE_CG_SINGLE;
    if (s == "direct") return RBMGPU_SOLVE_DIRECT;
    if (s == "cg-dynamic") {
        const double fs = 2. * n * n * (double)(m + t) - (double)n * (n + t);
        const double fd = 4. * n * (double)m * t - (double)t * (m + n);
        std::cout << "using CG method: " << (fd < fs ? "CG" : "CG_SINGLE") << std::endl;
        return fd < fs ? RBMGPU_SOLVE_CG : RBMGPU_SOLVE_CG_SINGLE;
    }
    throw std::runtime_error("optimizer.sr.method '" + s + "' is not supported by the GPU backend");
}
}  // namespace

gpu_solver::gpu_solver(rbmgpu_ctx* ctx, size_t n, size_t n_samples, size_t max_iterations, double rtol,
                       const std::string& method)
    : ctx_(ctx), n_(n), max_it_(std::min(2 * n, max_iterations)), rtol_(rtol) {
    if (max_it_ == 0) max_it_ = 2 * n;  // cg_solver.cpp:48-56
    if (rtol_ == 0.) rtol_ = 1e-16;
    method_ = solver_method(method, n, n_samples, max_it_);
}

void gpu_solver::solve(const std::vector<std::complex<double>>& b, std::vector<std::complex<double>>& x, double r1,
                       double r2, double rd) {
    if (x.size() != b.size()) x.assign(b.size(), 0.);
    check(ctx_,
          rbmgpu_sr_solve_ex(ctx_, method_, (const double*)b.data(), (double*)x.data(), r1, r2, rd,
                             (int64_t)max_it_, rtol_, &itn_, &rs_),
          "sr_solve");
}

void heun_plugin::apply(std::vector<std::complex<double>>& dw, double lr) {
    const size_t n = dw.size();
    std::vector<std::complex<double>> delta(n);
    for (size_t i = 0; i < n; i++) delta[i] = 0.5 * lr * dw[i];
    rbm_.update_weights_nc(delta);
    sampler_.sample();
    // the optimizer's gradient overwrites its own step, which dw refers to
    // (plugin.cpp:97 reads it back through the same vector)
    const std::vector<std::complex<double>>& g = opt_.gradient(false);
    for (size_t i = 0; i < n; i++) delta[i] = (delta[i] - 0.5 * lr * g[i]) / 6.;
    double x = 0;
    if (opt_.has_metric()) {
        // sum_k |delta^T O_k - delta^T <O>|^2 = S delta^H S0 delta, S0 the
        // unregularised S of the current samples (plugin.cpp:100-104)
        std::vector<std::complex<double>> y(n);
        rbmgpu_ctx* c = rbm_.ctx();
        check(c, rbmgpu_apply_S(c, (const double*)delta.data(), (double*)y.data(), 0., 0., 0.), "apply_S");
        std::complex<double> q = 0;
        for (size_t i = 0; i < n; i++) q += std::conj(delta[i]) * y[i];
        x = std::sqrt(std::max(q.real(), 0.) * (double)sampler_.get_my_n_samples()) / (double)n;
    } else {
        for (size_t i = 0; i < n; i++) x += std::norm(delta[i]);
        x = std::sqrt(x);
    }
    x = std::pow(eps_ / x, 1. / 3);
    for (size_t i = 0; i < n; i++) dw[i] *= (x - 0.5) * lr;
}

abstract_optimizer::abstract_optimizer(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s,
                                       const ini::config& cfg)
    : rbm_(rbm),
      sampler_(s),
      lr_(cfg.opt_lr, rbm.get_n_updates()),
      resample_(cfg.opt_resample),
      alpha1_(cfg.opt_resample_alpha1),
      alpha2_(cfg.opt_resample_alpha2),
      alpha3_(cfg.opt_resample_alpha3),
      dw_(rbm.get_n_params(), 0.),
      last_update_(rbm.get_n_params(), 0.) {}

void abstract_optimizer::log_energy(const char* std_key) {
    const auto& h = sampler_.energy();
    const double N = (double)rbm_.n_visible;
    log_values = {{"Energy", h.result[0].real() / N},
                  {std_key, h.get_stddev() / N},
                  {"AccetpanceRate", sampler_.get_acceptance_rate()}};
}

bool abstract_optimizer::optimize() {
    const double lr = lr_.get();
    const double N = (double)rbm_.n_visible;
    auto scale = [&](std::vector<std::complex<double>>& d) {
        if (plug_) plug_->apply(d, lr);
        else
            for (auto& v : d) v *= lr;
    };
    // resample heuristic (abstract_optimizer.cpp:83-129).  As there,
    // last_energy_ holds the previous epoch's energy per site while e is
    // the total energy of this one.
    if (resample_ && rbm_.get_n_updates() > 50) {
        std::complex<double> e = sampler_.energy().result.at(0);
        double d = std::sqrt(sampler_.energy().variance.at(0));
        int resample = 1;
        size_t rcount = 0;
        while (resample && last_energy_std_ != -1 && rcount < 3) {
            resample = 0;
            resample |= std::abs(std::real(last_energy_ - e) / N) >= alpha1_;
            resample |= (std::abs(std::imag(e)) >= alpha2_ * d) << 1;
            resample |= (d / last_energy_std_ >= alpha3_) << 2;
            resample |= (last_energy_std_ / d > alpha3_) << 3;
            if (resample) {
                std::cout << "resample " << resample << ", " << d << " " << last_energy_std_ << std::endl;
                rcount++;
                std::vector<std::complex<double>> undo(last_update_.size());
                for (size_t i = 0; i < undo.size(); i++) undo[i] = -last_update_[i];
                rbm_.update_weights_nc(undo);
                sampler_.sample();
                gradient(false);
                for (auto& v : dw_) v /= std::pow(2., (double)rcount);
                scale(dw_);
                rbm_.update_weights_nc(dw_);
                last_update_ = dw_;
                sampler_.sample();
                e = sampler_.energy().result.at(0);
                d = std::sqrt(sampler_.energy().variance.at(0));
            }
        }
        last_energy_ = e;
        last_energy_std_ = d;
        total_resamples_ += rcount;
    }
    gradient(true);
    last_energy_ = sampler_.energy().result.at(0) / N;
    if (!optim::step_ok(dw_.data(), dw_.size())) return false;  // abstract_optimizer.cpp:157
    scale(dw_);
    rbm_.update_weights(dw_);
    last_update_ = dw_;
    return true;
}

stochastic_reconfiguration::stochastic_reconfiguration(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s,
                                                       const ini::config& cfg)
    : abstract_optimizer(rbm, s, cfg),
      kp1_(cfg.opt_sr_reg1, rbm.get_n_updates()),
      kp2_(cfg.opt_sr_reg2, rbm.get_n_updates()),
      kp1d_(cfg.opt_sr_deltareg1, rbm.get_n_updates()),
      solver_(rbm.ctx(), rbm.get_n_params(), s.get_n_samples(), cfg.opt_sr_max_iterations, cfg.opt_sr_rtol,
              cfg.opt_sr_method),
      F_(rbm.get_n_params(), 0.),
      learned_(cfg.learn_hidden > 0) {
    metric_ = true;
}

std::vector<std::complex<double>>& stochastic_reconfiguration::gradient(bool log) {
    const auto& h = sampler_.energy();
    const auto& d = sampler_.derivative();
    const auto& dh = sampler_.energy_derivative();
    if (log) log_energy("Energy Stddev");
    double reg1 = kp1_.get(), reg2 = kp2_.get(), reg1delta = kp1d_.get();
    for (size_t p = 0; p < F_.size(); p++) F_[p] = dh.result[p] - h.result[0] * std::conj(d.result[p]);
    solver_.solve(F_, dw_, reg1, reg2, reg1delta);
    return dw_;
}

// the plain path of optimize() with the matrix-free CG runs as one device
// epoch: no plugin, no resample heuristic, no learned proposal to train
bool stochastic_reconfiguration::fused() const {
    return !plug_ && !resample_ && !learned_ && !sampler_.is_full() && solver_.method() == RBMGPU_SOLVE_CG;
}

bool stochastic_reconfiguration::train_step() {
    rbmgpu_sr_params p{};
    p.lr = lr_.get();  // abstract_optimizer.cpp:81, then SR.cpp:95-99
    p.reg1 = kp1_.get();
    p.reg2 = kp2_.get();
    p.reg1_delta = kp1d_.get();
    p.max_iter = (int64_t)solver_.max_iterations();
    p.rtol = solver_.rtol();
    p.graph = 1;
    rbmgpu_epoch_stats st{};
    rbmgpu_ctx* c = rbm_.ctx();
    check(c, rbmgpu_train_epoch(c, sampler_.take_epoch(), &p, &st), "train_epoch");
    sampler_.set_epoch_stats(st);
    log_energy("Energy Stddev");
    last_energy_ = sampler_.energy().result.at(0) / (double)rbm_.n_visible;
    if (!st.updated) return false;
    rbm_.note_device_update();
    return true;
}

gradient_descent::gradient_descent(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s,
                                   const ini::config& cfg)
    : abstract_optimizer(rbm, s, cfg), real_factor_(cfg.opt_sgd_real_factor) {}

std::vector<std::complex<double>>& gradient_descent::gradient(bool log) {
    const auto& h = sampler_.energy();
    if (log) log_energy("EnergyStddev");
    optim::sgd_gradient(sampler_.energy_derivative().result.data(), sampler_.derivative().result.data(),
                        h.result[0], real_factor_, dw_.data(), dw_.size());
    return dw_;
}

namespace {
class host_plugin : public base_plugin {
   public:
    explicit host_plugin(std::unique_ptr<optim::plugin> p) : p_(std::move(p)) {}
    void apply(std::vector<std::complex<double>>& dw, double lr) override { p_->apply(dw.data(), dw.size(), lr); }

   private:
    std::unique_ptr<optim::plugin> p_;
};
}  // namespace

std::unique_ptr<abstract_optimizer> make_optimizer(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s,
                                                   const ini::config& cfg) {
    std::unique_ptr<abstract_optimizer> opt;
    if (cfg.opt_type == "SR") opt.reset(new stochastic_reconfiguration(rbm, s, cfg));
    else if (cfg.opt_type == "SGD") opt.reset(new gradient_descent(rbm, s, cfg));
    else throw std::runtime_error("optimizer.type '" + cfg.opt_type + "' is not supported by the GPU backend");
    const size_t P = rbm.get_n_params();
    if (cfg.opt_plugin == "adam") {
        opt->set_plugin(std::unique_ptr<base_plugin>(new host_plugin(std::unique_ptr<optim::plugin>(
            new optim::adam(P, cfg.opt_adam_beta1, cfg.opt_adam_beta2, cfg.opt_adam_eps)))));
    } else if (cfg.opt_plugin == "momentum") {
        opt->set_plugin(std::unique_ptr<base_plugin>(new host_plugin(
            std::unique_ptr<optim::plugin>(new optim::momentum(P, cfg.opt_mom_alpha, cfg.opt_mom_dialup)))));
    } else if (cfg.opt_plugin == "heun") {
        opt->set_plugin(std::unique_ptr<base_plugin>(new heun_plugin(*opt, rbm, s, cfg.opt_heun_eps)));
    } else if (!cfg.opt_plugin.empty()) {
        throw std::runtime_error("optimizer.plugin '" + cfg.opt_plugin + "' is unknown");
    }
    return opt;
}

}  // namespace optimizer
