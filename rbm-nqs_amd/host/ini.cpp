// ini.cpp -- see ini.hpp.
#include "ini.hpp"

#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace ini {

namespace {

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// comma / blank separated list (ini.cpp:338-412 split at ',')
std::vector<std::string> split(const std::string& s) {
    std::vector<std::string> out;
    std::string tok;
    for (char c : s) {
        if (c == ',' || c == ' ' || c == '\t') {
            if (!tok.empty()) out.push_back(tok);
            tok.clear();
        } else {
            tok += c;
        }
    }
    if (!tok.empty()) out.push_back(tok);
    return out;
}

std::vector<double> dlist(const std::string& s) {
    std::vector<double> v;
    for (auto& t : split(s)) v.push_back(std::stod(t));
    if (v.empty()) throw std::invalid_argument("empty list");
    return v;
}

decay_t decay(const std::string& s) {
    auto v = dlist(s);
    if (v.size() == 3) return {v[0], v[1], v[2]};
    return {v[0], v[0], 1.};
}

bool boolean(const std::string& s) {
    if (s == "1" || s == "true" || s == "on" || s == "yes") return true;
    if (s == "0" || s == "false" || s == "off" || s == "no") return false;
    throw std::invalid_argument("bad boolean '" + s + "'");
}

using setter = std::function<void(config&, const std::string&)>;

const std::map<std::string, setter>& table() {
    static const std::map<std::string, setter> t = {
        {"seed", [](config& c, const std::string& v) { c.seed = std::stoull(v); }},
        {"deterministic", [](config& c, const std::string& v) { c.deterministic = boolean(v); }},
        {"name", [](config& c, const std::string& v) { c.name = v; }},
        {"n_threads", [](config& c, const std::string& v) { c.n_threads = std::stoul(v); }},
        {"seed_search", [](config& c, const std::string& v) { c.seed_search = std::stoi(v); }},
        {"seed_search_epochs", [](config& c, const std::string& v) { c.seed_search_epochs = std::stoul(v); }},
        {"n_epochs", [](config& c, const std::string& v) { c.n_epochs = std::stoul(v); }},
        {"model.type", [](config& c, const std::string& v) {
             if (v != "kitaev" && v != "kitaevS3" && v != "isingS3" && v != "toric")
                 throw std::runtime_error("Model Type '" + v + "' not available!");
             c.model = v;
         }},
        {"model.n_cells", [](config& c, const std::string& v) { c.n_cells = std::stoul(v); }},
        {"model.n_cells_b", [](config& c, const std::string& v) { c.n_cells_b = std::stoi(v); }},
        {"model.J", [](config& c, const std::string& v) { c.J = dlist(v); }},
        {"model.h", [](config& c, const std::string& v) { c.h = std::stod(v); }},
        {"model.helper_strength", [](config& c, const std::string& v) { c.helper_strength = std::stod(v); }},
        {"model.symmetry", [](config& c, const std::string& v) { c.symmetry = dlist(v); }},
        {"model.lattice_type", [](config& c, const std::string& v) { c.lattice_type = v; }},
        {"rbm.type", [](config& c, const std::string& v) {
             if (v != "basic" && v != "symmetry" && v != "pfaffian" && v != "file")
                 throw std::runtime_error("RBM Type '" + v + "' not available!");
             c.rbm = v;
         }},
        {"rbm.alpha", [](config& c, const std::string& v) { c.alpha = std::stoul(v); }},
        {"rbm.force", [](config& c, const std::string& v) { c.rbm_force = boolean(v); }},
        {"rbm.weights", [](config& c, const std::string& v) { c.rbm_weights = std::stod(v); }},
        {"rbm.weights_imag", [](config& c, const std::string& v) { c.rbm_weights_imag = std::stod(v); }},
        {"rbm.weights_type", [](config& c, const std::string& v) { c.rbm_weights_init_type = v; }},
        {"rbm.correlators", [](config& c, const std::string& v) { c.rbm_correlators = std::stoul(v); }},
        {"rbm.pop_mode", [](config& c, const std::string& v) { c.rbm_pop_mode = std::stoul(v); }},
        {"rbm.cosh_mode", [](config& c, const std::string& v) { c.rbm_cosh_mode = std::stoul(v); }},
        {"rbm.pfaffian", [](config& c, const std::string& v) { c.rbm_pfaffian = boolean(v); }},
        {"rbm.pfaffian.symmetry", [](config& c, const std::string& v) { c.rbm_pfaffian_symmetry = dlist(v); }},
        {"rbm.pfaffian.weights", [](config& c, const std::string& v) { c.rbm_pfaffian_weights = std::stod(v); }},
        {"rbm.pfaffian.normalize", [](config& c, const std::string& v) { c.rbm_pfaffian_normalize = boolean(v); }},
        {"rbm.file.name", [](config& c, const std::string& v) { c.rbm_file_name = v; }},
        {"rbm.pfaffian.load", [](config& c, const std::string& v) { c.rbm_pfaffian_load = v; }},
        {"rbm.pfaffian.no_updating", [](config& c, const std::string& v) { c.rbm_pfaffian_no_updating = boolean(v); }},
        {"sampler.type", [](config& c, const std::string& v) {
             if (v != "full" && v != "exact" && v != "metropolis")
                 throw std::runtime_error("Sampler Type '" + v + "' not available!");
             c.sa_type = v;
         }},
        {"sampler.n_samples", [](config& c, const std::string& v) { c.sa_n_samples = std::stoul(v); }},
        {"sampler.n_samples_per_chain",
         [](config& c, const std::string& v) { c.sa_metropolis_samples_per_chain = std::stoul(v); }},
        {"sampler.eval_samples", [](config& c, const std::string& v) { c.sa_eval_samples = std::stoul(v); }},
        {"sampler.full.n_parallel_bits", [](config& c, const std::string& v) { c.sa_full_n_parallel_bits = std::stoul(v); }},
        {"sampler.metropolis.n_chains", [](config& c, const std::string& v) { c.sa_metropolis_n_chains = std::stoul(v); }},
        {"sampler.metropolis.n_warmup_steps",
         [](config& c, const std::string& v) { c.sa_metropolis_n_warmup_steps = std::stoul(v); }},
        {"sampler.metropolis.n_steps_per_sample",
         [](config& c, const std::string& v) { c.sa_metropolis_n_steps_per_sample = std::stoul(v); }},
        {"sampler.metropolis.bond_flips", [](config& c, const std::string& v) { c.sa_metropolis_bond_flips = std::stod(v); }},
        {"sampler.pfaffian_refresh", [](config& c, const std::string& v) { c.sa_pfaffian_refresh = std::stoi(v); }},
        {"sampler.lut_exchange", [](config& c, const std::string& v) { c.sa_lut_exchange = std::stoi(v); }},
        {"optimizer.type", [](config& c, const std::string& v) {
             if (v != "SR" && v != "SGD") throw std::runtime_error("Optimizer Type '" + v + "' not available!");
             c.opt_type = v;
         }},
        {"optimizer.learning_rate", [](config& c, const std::string& v) { c.opt_lr = decay(v); }},
        {"optimizer.sr.reg1", [](config& c, const std::string& v) { c.opt_sr_reg1 = decay(v); }},
        {"optimizer.sr.reg2", [](config& c, const std::string& v) { c.opt_sr_reg2 = decay(v); }},
        {"optimizer.sr.deltareg1", [](config& c, const std::string& v) { c.opt_sr_deltareg1 = decay(v); }},
        {"optimizer.sr.method", [](config& c, const std::string& v) { c.opt_sr_method = v; }},
        {"optimizer.sr.iterative", [](config& c, const std::string& v) { c.opt_sr_iterative = boolean(v); }},
        {"optimizer.sr.max_iterations", [](config& c, const std::string& v) { c.opt_sr_max_iterations = std::stoul(v); }},
        {"optimizer.sr.rtol", [](config& c, const std::string& v) { c.opt_sr_rtol = std::stod(v); }},
        {"optimizer.sgd.real_factor", [](config& c, const std::string& v) { c.opt_sgd_real_factor = std::stod(v); }},
        {"optimizer.plugin", [](config& c, const std::string& v) { c.opt_plugin = v; }},
        {"optimizer.adam.beta1", [](config& c, const std::string& v) { c.opt_adam_beta1 = std::stod(v); }},
        {"optimizer.adam.beta2", [](config& c, const std::string& v) { c.opt_adam_beta2 = std::stod(v); }},
        {"optimizer.adam.eps", [](config& c, const std::string& v) { c.opt_adam_eps = std::stod(v); }},
        {"optimizer.mom.alpha", [](config& c, const std::string& v) { c.opt_mom_alpha = std::stod(v); }},
        {"optimizer.mom.dialup", [](config& c, const std::string& v) { c.opt_mom_dialup = std::stod(v); }},
        {"optimizer.heun.eps", [](config& c, const std::string& v) { c.opt_heun_eps = std::stod(v); }},
        {"optimizer.resample", [](config& c, const std::string& v) { c.opt_resample = boolean(v); }},
        {"optimizer.resample.alpha1", [](config& c, const std::string& v) { c.opt_resample_alpha1 = std::stod(v); }},
        {"optimizer.resample.alpha2", [](config& c, const std::string& v) { c.opt_resample_alpha2 = std::stod(v); }},
        {"optimizer.resample.alpha3", [](config& c, const std::string& v) { c.opt_resample_alpha3 = std::stod(v); }},
        {"gpu.precision", [](config& c, const std::string& v) {
             int p = std::stoi(v);
             if (p != 32 && p != 64) throw std::invalid_argument("gpu.precision must be 32 or 64");
             c.gpu_precision = p;
         }},
        {"gpu.lanes_per_chain", [](config& c, const std::string& v) { c.gpu_lanes = std::stoi(v); }},
        {"gpu.keep_o", [](config& c, const std::string& v) { c.gpu_keep_o = boolean(v); }},
        {"gpu.auto_chains", [](config& c, const std::string& v) { c.gpu_auto_chains = boolean(v); }},
        {"gpu.explicit_s", [](config& c, const std::string& v) {
             c.gpu_explicit_s = v == "auto" ? -1 : (boolean(v) ? 1 : 0);
         }},
        {"learn.hidden", [](config& c, const std::string& v) { c.learn_hidden = std::stoi(v); }},
        {"learn.fraction", [](config& c, const std::string& v) { c.learn_fraction = std::stod(v); }},
        {"learn.lr", [](config& c, const std::string& v) { c.learn_lr = std::stod(v); }},
        {"learn.entropy", [](config& c, const std::string& v) { c.learn_entropy = std::stod(v); }},
        {"learn.batch", [](config& c, const std::string& v) { c.learn_batch = std::stoll(v); }},
    };
    return t;
}

// switches (po::bool_switch) and their short forms
const std::map<std::string, bool config::*>& switches() {
    static const std::map<std::string, bool config::*> s = {
        {"train", &config::train},           {"evaluate", &config::evaluate},
        {"store_state", &config::store_state}, {"store_samples", &config::store_samples},
        {"print_bonds", &config::print_bonds}, {"print_hex", &config::print_hex},
        {"exact_energy", &config::exact_energy}, {"noprogress", &config::noprogress},
        {"rbm.force", &config::rbm_force},   {"rbm.pfaffian.no_updating", &config::rbm_pfaffian_no_updating},
    };
    return s;
}
const std::map<char, std::string>& shorts() {
    static const std::map<char, std::string> s = {{'i', "infile"}, {'n', "name"}, {'c', "model.n_cells"},
                                                  {'t', "n_threads"}, {'P', "noprogress"}, {'f', "rbm.force"},
                                                  {'l', "optimizer.learning_rate"}, {'e', "n_epochs"},
                                                  {'h', "help"}};
    return s;
}

}  // namespace

void set_option(config& cfg, const std::string& key, const std::string& value) {
    auto it = table().find(key);
    if (it == table().end()) {
        auto sw = switches().find(key);
        if (sw != switches().end()) {
            cfg.*(sw->second) = boolean(value);
            return;
        }
        throw std::runtime_error("unrecognised option '" + key + "'");
    }
    it->second(cfg, value);
}

void parse_ini_text(const std::string& text, config& cfg) {
    std::istringstream in(text);
    std::string line, section;
    while (std::getline(in, line)) {
        size_t c = line.find('#');
        if (c != std::string::npos) line = line.substr(0, c);
        line = trim(line);
        if (line.empty() || line[0] == ';') continue;
        if (line[0] == '[') {
            size_t e = line.find(']');
            if (e == std::string::npos) throw std::runtime_error("bad section line '" + line + "'");
            section = trim(line.substr(1, e - 1));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) throw std::runtime_error("bad ini line '" + line + "'");
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (!section.empty()) key = section + "." + key;
        set_option(cfg, key, val);
    }
}

void finalize(config& cfg, bool name_given) {
    if (!name_given) cfg.name = "n" + std::to_string(cfg.n_cells) + "_a" + std::to_string(cfg.alpha);
    if (cfg.sa_metropolis_samples_per_chain)
        cfg.sa_n_samples = cfg.sa_metropolis_n_chains * cfg.sa_metropolis_samples_per_chain;
}

int load(int argc, char** argv, config& cfg) {
    std::vector<std::pair<std::string, std::string>> cli;
    bool help = false;
    try {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i], key, val;
            bool has_val = false;
            if (a.rfind("--", 0) == 0) {
                key = a.substr(2);
                size_t eq = key.find('=');
                if (eq != std::string::npos) {
                    val = key.substr(eq + 1);
                    key = key.substr(0, eq);
                    has_val = true;
                }
            } else if (a.size() >= 2 && a[0] == '-') {
                auto s = shorts().find(a[1]);
                if (s == shorts().end()) throw std::runtime_error("unrecognised option '" + a + "'");
                key = s->second;
                if (a.size() > 2) {
                    val = a.substr(2);
                    has_val = true;
                }
            } else {
                throw std::runtime_error("unexpected argument '" + a + "'");
            }
            if (key == "help") {
                help = true;
                continue;
            }
            if (switches().count(key) && !has_val) {
                cli.push_back({key, "1"});
                continue;
            }
            if (!has_val) {
                if (i + 1 >= argc) throw std::invalid_argument("missing value for '" + key + "'");
                val = argv[++i];
            }
            if (key == "infile") {
                cfg.ini_file = val;
                continue;
            }
            cli.push_back({key, val});
        }
        if (help) {
            std::cout << "rbm: options are the keys of the reference's .ini files, e.g. --model.n_cells=4\n";
            return 1;
        }
        bool name_given = false;
        if (!cfg.ini_file.empty()) {
            std::ifstream f(cfg.ini_file);
            if (!f.good()) throw std::runtime_error("INI File '" + cfg.ini_file + "' does not exist!");
            std::stringstream ss;
            ss << f.rdbuf();
            config from_file = cfg;
            parse_ini_text(ss.str(), from_file);
            name_given = from_file.name != cfg.name;
            cfg = from_file;
        }
        for (auto& kv : cli) {
            if (kv.first == "name") name_given = true;
            set_option(cfg, kv.first, kv.second);
        }
        finalize(cfg, name_given);
    } catch (const std::invalid_argument& e) {
        std::cerr << "Invalid argument: " << e.what() << std::endl;
        return -2;
    } catch (const std::exception& e) {
        std::cerr << e.what();
        return -1;
    }
    return 0;
}

}  // namespace ini
