// ini.hpp -- run configuration with the reference's option names and
// defaults (src/tools/ini.cpp:38-122, option table 154-234).  Values come
// from a .ini file ("[section]" + "key = value" -> "section.key") and from
// command-line options "--section.key=value" / "--section.key value", the
// command line taking precedence like boost::program_options does.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ini {

struct decay_t {
    double initial, min, decay = 1.;
};

struct config {
    // program
    uint64_t seed = 421390484ULL;
    size_t n_threads = 4;
    std::string name;
    std::string ini_file;
    bool train = false, evaluate = false, store_state = false, store_samples = false;
    bool noprogress = false, print_bonds = false, print_hex = false, deterministic = true;
    bool exact_energy = false;
    int seed_search = 0;
    size_t seed_search_epochs = 1000;
    // model
    std::string model = "kitaev";
    size_t n_cells = 2;
    int n_cells_b = -1;
    std::vector<double> J{-1.};
    double h = 0., helper_strength = 0.;
    std::vector<double> symmetry{0.5};
    std::string lattice_type;
    // rbm
    std::string rbm = "basic";
    size_t alpha = 1;
    bool rbm_force = false;
    size_t rbm_pop_mode = 0, rbm_cosh_mode = 2;
    double rbm_weights = 1e-4, rbm_weights_imag = -1;
    std::string rbm_weights_init_type;
    size_t rbm_correlators = 0;
    bool rbm_pfaffian = false;
    std::vector<double> rbm_pfaffian_symmetry{0.5};
    bool rbm_pfaffian_normalize = false;
    double rbm_pfaffian_weights = 0.01;
    bool rbm_pfaffian_no_updating = false;
    std::string rbm_file_name, rbm_pfaffian_load;
    // sampler
    std::string sa_type = "metropolis";
    size_t sa_n_samples = 1000, sa_eval_samples = 0, sa_metropolis_samples_per_chain = 0;
    size_t sa_metropolis_n_chains = 4, sa_metropolis_n_warmup_steps = 100;
    size_t sa_metropolis_n_steps_per_sample = 1, sa_full_n_parallel_bits = 2;
    double sa_metropolis_bond_flips = 0.5;
    int sa_pfaffian_refresh = 0, sa_lut_exchange = 0;
    // optimizer
    std::string opt_type = "SR", opt_plugin;
    decay_t opt_lr{1e-2, 1e-2, 1.};
    decay_t opt_sr_reg1{1, 1e-4, 0.9}, opt_sr_reg2{1e-3, 1e-6, 0.9}, opt_sr_deltareg1{1e-2, 1e-2, 1.};
    double opt_sgd_real_factor = 1., opt_adam_beta1 = 0.9, opt_adam_beta2 = 0.999, opt_adam_eps = 1e-8;
    double opt_mom_alpha = 0.9, opt_mom_dialup = 1.;
    std::string opt_sr_method = "cg";
    bool opt_sr_iterative = true;
    size_t opt_sr_max_iterations = 0;
    double opt_sr_rtol = 0.;
    double opt_heun_eps = 1e-3;
    bool opt_resample = false;
    double opt_resample_alpha1 = 0.1, opt_resample_alpha2 = 5, opt_resample_alpha3 = 10;
    // train
    size_t n_epochs = 4000;
    // GPU backend (new keys; the reference has no GPU)
    int gpu_precision = 32;
    int gpu_lanes = 0;
    // materialise the O matrix even where the structured S.x applies
    bool gpu_keep_o = false;
    bool gpu_auto_chains = false;  // the autotuner picks the chain count
    int gpu_explicit_s = -1;       // SR epoch CG on the explicit S: 1 on, 0 off, -1 cost model
    // learned cluster proposal (new keys; the reference has none)
    int learn_hidden = 0;          // hidden width of the proposal network, 0: off
    double learn_fraction = 0.5;   // probability of a learned move per slot
    double learn_lr = 1e-3;        // Adam learning rate
    double learn_entropy = 0.01;   // entropy weight of the objective
    int64_t learn_batch = 16384;   // records per training step
};

// Parse .ini text into cfg (unknown keys are an error, like
// po::parse_config_file).  Throws std::runtime_error.
void parse_ini_text(const std::string& text, config& cfg);
// Parse argv (reads -i/--infile first, then the file, then the command
// line).  Returns 0, or the reference's codes (-1 unknown option,
// -2 invalid argument, 1 for --help).
int load(int argc, char** argv, config& cfg);
// set one dotted key from its string value
void set_option(config& cfg, const std::string& key, const std::string& value);
// post-processing of ini.cpp:252-262 (default name, samples per chain)
void finalize(config& cfg, bool name_given);

}  // namespace ini
