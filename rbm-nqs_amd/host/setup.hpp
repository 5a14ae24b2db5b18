// setup.hpp -- builds the model and the GPU context from a configuration,
// like the factories init_model / init_machine / init_sampler of
// src/main.cpp:125-239, and owns the tables the rbmgpu_config points to.
#pragma once
#include <complex>
#include <memory>
#include <random>
#include <vector>

#include "../../include/rbmgpu.h"
#include "ini.hpp"
#include "model.hpp"

namespace rbmhost {

struct machine_layout {
    int n_spins = 0, n_vb = 0, n_alpha = 0, n_symm = 1, pf_rows = 0;
    std::vector<int32_t> perm, basis, pf_bs;
    int64_t n_params() const { return (int64_t)n_vb + n_alpha + (int64_t)n_alpha * n_spins + 4 * pf_rows; }
    int64_t n_neural() const { return (int64_t)n_vb + n_alpha + (int64_t)n_alpha * n_spins; }
};

// observables: group 0 is the Hamiltonian, then optional extra groups
struct term_table {
    std::vector<int32_t> k, sites, group;
    std::vector<double> mat;
    int n_groups = 1;
    void add(const model::local_op& op, int g);
};

class setup {
   public:
    explicit setup(const ini::config& cfg);
    const ini::config& cfg() const { return cfg_; }
    model::abstract_model& model() { return *model_; }
    const machine_layout& layout() const { return layout_; }
    const term_table& terms() const { return terms_; }
    const std::vector<int32_t>& prop_bonds() const { return prop_bonds_; }
    // cluster catalogue of the learned proposal: N x kmax, row i = the
    // first lattice loop (Kitaev hexagon, toric vertex star) through site i,
    // starting at i
    const std::vector<int32_t>& clusters() const { return clusters_; }
    int cluster_kmax() const { return cluster_kmax_; }

    // Hamiltonian only (training), or Hamiltonian without helper terms plus
    // one group per hexagon plaquette (--evaluate on Kitaev, main.cpp:640-680)
    void use_evaluation_observables(bool hexagons);
    // number of samples (set_n_samples of the sampler, metropolis_sampler.cpp:289-293)
    void set_n_samples(size_t n);
    size_t n_samples() const { return n_samples_; }
    size_t warmup() const { return warmup_; }

    rbmgpu_config gpu_config() const;
    // rbm_base::initialize_weights with the reference's RNG draws
    // (rbm_base.cpp:58-113) then the Pfaffian weights (pfaffian.cpp:139-155)
    std::vector<std::complex<double>> initial_weights(std::mt19937& rng) const;

   private:
    ini::config cfg_;
    std::unique_ptr<model::abstract_model> model_;
    machine_layout layout_;
    term_table terms_;
    std::vector<int32_t> prop_bonds_;
    std::vector<int32_t> clusters_;
    int cluster_kmax_ = 0;
    size_t n_samples_, warmup_;
};

}  // namespace rbmhost
