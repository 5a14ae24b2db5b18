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
    push();
}

void gpu_rbm::initialize_weights(std::mt19937& rng, double std_dev, double std_dev_imag, double pf_std) {
    n_updates_ = 0;
    // the draw order of rbm_base::initialize_weights, then the Pfaffian
    std::normal_distribution<double> real_dist{0, std_dev},
        imag_dist{0, std_dev_imag < 0 ? std_dev : std_dev_imag};
    const auto& L = setup_.layout();
    std::vector<cplx> p(n_params_, 0.);
    const int N = L.n_spins, nvb = L.n_vb, na = L.n_alpha;
    for (int i = 0; i < nvb; i++) p[i] = cplx(real_dist(rng), imag_dist(rng));
    for (int i = 0; i < na; i++) {
        p[nvb + i] = cplx(real_dist(rng), imag_dist(rng));
        for (int j = 0; j < N; j++) p[nvb + na + j + (int64_t)i * N] = cplx(real_dist(rng), imag_dist(rng));
    }
    if (has_pfaffian()) {
        std::normal_distribution<double> dist{0, pf_std};
        for (int64_t i = n_neural_; i < n_params_; i++) p[i] = cplx(dist(rng), dist(rng));
    }
    set_params(p);
}

void gpu_rbm::update_weights(const std::vector<cplx>& dw) {
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
    : rbm_(rbm), n_samples_(s.n_samples()), groups_(s.terms().n_groups) {}

void gpu_metropolis_sampler::sample() {
    rbmgpu_ctx* c = rbm_.ctx();
    check(c, rbmgpu_sample_epoch(c, epoch_++), "sample_epoch");
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

void gpu_cg_solver::solve(const std::vector<std::complex<double>>& b, std::vector<std::complex<double>>& x,
                          double r1, double r2, double rd) {
    if (x.size() != b.size()) x.assign(b.size(), 0.);
    check(ctx_,
          rbmgpu_sr_solve(ctx_, (const double*)b.data(), (double*)x.data(), r1, r2, rd, (int64_t)max_it_, rtol_,
                          &itn_, &rs_),
          "sr_solve");
}

stochastic_reconfiguration::stochastic_reconfiguration(machine::gpu_rbm& rbm, sampler::gpu_metropolis_sampler& s,
                                                       const ini::config& cfg)
    : rbm_(rbm),
      sampler_(s),
      lr_(cfg.opt_lr, rbm.get_n_updates()),
      kp1_(cfg.opt_sr_reg1, rbm.get_n_updates()),
      kp2_(cfg.opt_sr_reg2, rbm.get_n_updates()),
      kp1d_(cfg.opt_sr_deltareg1, rbm.get_n_updates()),
      solver_(rbm.ctx(), rbm.get_n_params(), cfg.opt_sr_max_iterations, cfg.opt_sr_rtol),
      dw_(rbm.get_n_params(), 0.),
      F_(rbm.get_n_params(), 0.) {
    if (cfg.opt_sr_method != "cg" && cfg.opt_sr_method != "minres" && cfg.opt_sr_method != "cg-single" &&
        cfg.opt_sr_method != "cg-dynamic")
        throw std::runtime_error("optimizer.sr.method '" + cfg.opt_sr_method + "' is not supported by the GPU backend");
    if (cfg.opt_type != "SR") throw std::runtime_error("only the SR optimizer runs on the GPU backend");
    if (!cfg.opt_plugin.empty()) throw std::runtime_error("optimizer plugins are not supported by the GPU backend");
}

std::vector<std::complex<double>>& stochastic_reconfiguration::gradient(bool log) {
    const auto& h = sampler_.energy();
    const auto& d = sampler_.derivative();
    const auto& dh = sampler_.energy_derivative();
    const double N = (double)rbm_.n_visible;
    if (log) {
        log_values = {{"Energy", h.result[0].real() / N},
                      {"Energy Stddev", h.get_stddev() / N},
                      {"AccetpanceRate", sampler_.get_acceptance_rate()}};
    }
    double reg1 = kp1_.get(), reg2 = kp2_.get(), reg1delta = kp1d_.get();
    for (size_t p = 0; p < F_.size(); p++) F_[p] = dh.result[p] - h.result[0] * std::conj(d.result[p]);
    solver_.solve(F_, dw_, reg1, reg2, reg1delta);
    return dw_;
}

bool stochastic_reconfiguration::optimize() {
    double lr = lr_.get();
    gradient(true);
    if (!optim::step_ok(dw_.data(), dw_.size())) return false;  // abstract_optimizer.cpp:157
    for (auto& v : dw_) v *= lr;
    rbm_.update_weights(dw_);
    return true;
}

double stochastic_reconfiguration::last_energy_per_site() const {
    return get_current_energy() / (double)rbm_.n_visible;
}

}  // namespace optimizer
