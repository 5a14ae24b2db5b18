// main.cpp -- the `rbm` driver with the reference's command line and flow
// (src/main.cpp:381-717): --train runs SR epochs on the GPU, --evaluate
// samples the energy (and the Kitaev hexagon plaquettes) with binning
// statistics, --print_bonds / --print_hex / --exact_energy print model
// tables.  Checkpoints are the reference's .rbm files, the per-epoch log is
// its <name>.log CSV.
#include <sys/ioctl.h>
#include <unistd.h>

#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <random>
#include <sstream>

#include "adapters.hpp"
#include "ini.hpp"
#include "setup.hpp"

namespace {

// <name>.log: "# header, ..." then one CSV line per epoch (logger.cpp:31-74)
class epoch_log {
   public:
    epoch_log(const std::string& name, bool force) {
        std::string fn = name + ".log";
        std::ifstream probe{fn};
        std::cout << "Logging to '" << fn << "'!" << std::endl;
        out_.open(fn, (probe.good() && !force) ? std::ios::app : std::ios::trunc);
    }
    void line(const std::vector<std::pair<std::string, double>>& vals) {
        if (first_) {
            std::string h;
            for (auto& kv : vals) h += kv.first + ", ";
            if (h.size() > 2) out_ << "# " << h.substr(0, h.size() - 2) << "\n";
            first_ = false;
        }
        std::ostringstream s;
        s.precision(16);
        for (size_t i = 0; i < vals.size(); i++) s << (i ? ", " : "") << vals[i].second;
        out_ << s.str() << std::endl;
    }

   private:
    std::ofstream out_;
    bool first_ = true;
};

void progress_bar(size_t i, size_t n, double energy, char st) {
    int width = 100;
    winsize ws{};
    if (ioctl(STDOUT_FILENO, TIOCGWINSZ, &ws) == 0 && ws.ws_col > 0 && ws.ws_col <= 500) width = ws.ws_col;
    double prog = n ? (double)i / n : 1.0;
    int plen = std::max(10, width - (2 * (int)std::log10(std::max<size_t>(n, 1)) + 37));
    int p = (int)(plen * prog + 0.5);
    std::cout << "\rEpochs: (" << i << "/" << n << ") " << st << " [" << std::string(p, '#')
              << std::string(plen - p, ' ') << "] Energy: " << std::setprecision(5) << energy << std::flush;
}

int special_cases(const ini::config& cfg) {
    if (cfg.print_bonds) {  // main.cpp:333-351
        rbmhost::setup s(cfg);
        if (cfg.model == "toric") {
            auto& lat = static_cast<lattice::toric_lattice&>(s.model().get_lattice());
            for (auto& p : lat.construct_plaqs()) {
                for (auto i : p.idxs) std::cout << i << ",";
                std::cout << p.type << "\n";
            }
        } else {
            for (auto& b : s.model().get_lattice().get_bonds()) std::cout << b.a << "," << b.b << "," << b.type << "\n";
        }
        return 1;
    }
    if (cfg.print_hex && cfg.model == "kitaev") {
        rbmhost::setup s(cfg);
        for (auto& h : static_cast<model::kitaev&>(s.model()).honeycomb_lattice().get_hexagons()) {
            for (auto x : h) std::cout << x << ", ";
            std::cout << "\n";
        }
        return 1;
    }
    if (cfg.exact_energy && cfg.model == "kitaev") {
        rbmhost::setup s(cfg);
        std::cout.precision(16);
        std::cout << "Exact Energy: " << static_cast<model::kitaev&>(s.model()).exact_energy() << std::endl;
        return 1;
    }
    return 0;
}

// --store_state (main.cpp:289-302, full_sampler.cpp:150-160): psi of every
// basis state, normalised, indexed by the state's number (bit i <=> sigma_i =
// -1), written as an eigen_fstream matrix (2^N x 1 complex) to <name>.state
void store_state(machine::gpu_rbm& rbm, const rbmhost::setup& s, const ini::config& cfg) {
    std::cout << "Storing State..." << std::endl;
    rbmgpu_ctx* c = rbm.ctx();
    if (rbmgpu_full_epoch(c, (int32_t)cfg.sa_full_n_parallel_bits) != 0)
        throw std::runtime_error(rbmgpu_last_error(c));
    int64_t S = 0;
    rbmgpu_n_samples(c, &S);
    std::vector<std::complex<double>> lp(S), vec(S);
    std::vector<uint64_t> st(S * ((rbm.n_visible + 63) / 64));
    if (rbmgpu_full_weights(c, (double*)lp.data(), nullptr) != 0 ||
        rbmgpu_get_samples(c, st.data(), nullptr, nullptr, nullptr) != 0)
        throw std::runtime_error(rbmgpu_last_error(c));
    double mx = -INFINITY, tot = 0;
    for (auto& v : lp) mx = std::max(mx, v.real());
    const size_t nw = (rbm.n_visible + 63) / 64;
    for (int64_t k = 0; k < S; k++) {
        const std::complex<double> psi = std::exp(lp[k] - mx);
        vec[st[k * nw]] = psi;
        tot += std::norm(psi);
    }
    for (auto& v : vec) v /= std::sqrt(tot);
    std::ofstream f{cfg.name + ".state", std::ios::binary};
    const uint64_t rows = (uint64_t)S, cols = 1;
    f.write((const char*)&rows, 8);
    f.write((const char*)&cols, 8);
    f.write((const char*)vec.data(), vec.size() * sizeof(std::complex<double>));
    std::cout << "State stored to '" << cfg.name << ".state'" << std::endl;
    (void)s;
}

}  // namespace

int main(int argc, char** argv) {
    ini::config cfg;
    int rc = ini::load(argc, argv, cfg);
    if (rc != 0) return rc;
    try {
        if (special_cases(cfg)) return 0;
        if (cfg.sa_type != "metropolis" && cfg.sa_type != "full")
            throw std::runtime_error("sampler.type '" + cfg.sa_type + "' is not supported by the GPU backend");
        std::cout << "Starting '" << cfg.name << "'!" << std::endl;
        std::unique_ptr<std::mt19937> rng;
        if (cfg.deterministic) {
            rng.reset(new std::mt19937((std::mt19937::result_type)cfg.seed));
        } else {
            std::random_device rd;
            std::seed_seq seq{rd(), rd(), rd(), rd(), rd(), rd(), rd(), rd()};
            rng.reset(new std::mt19937(seq));
        }
        std::cout << "Seed: " << cfg.seed << std::endl;
        rbmhost::setup setup(cfg);
        machine::gpu_rbm rbm(setup);
        if (cfg.rbm_force || !rbm.load(cfg.name)) rbm.initialize_weights(*rng);
        rc = 0;
        if (cfg.store_state) {  // a special case of main.cpp:374-377: store and stop
            store_state(rbm, setup, cfg);
            return 0;
        }
        if (cfg.train) {
            epoch_log log(cfg.name, cfg.rbm_force);
            sampler::gpu_metropolis_sampler sampler(rbm, setup);
            auto opt = optimizer::make_optimizer(rbm, sampler, cfg);
            std::cout << "Number of parameters: " << rbm.get_n_params() << std::endl;
            for (size_t i = rbm.get_n_updates(); i < cfg.n_epochs; i++) {
                if (i > 0 && i % 100 == 0) rbm.save(cfg.name, true);
                bool ok;
                if (opt->fused()) {
                    // sampling and the SR step as one device epoch (one graph replay)
                    ok = opt->train_step();
                } else {
                    sampler.sample();
                    if (!cfg.noprogress) progress_bar(i + 1, cfg.n_epochs, opt->last_energy_per_site(), 'O');
                    ok = opt->optimize();
                }
                if (!ok) {
                    std::cerr << "\nFATAL: Massive step detected." << std::endl;
                    rc = 55;
                    break;
                }
                log.line(opt->log_values);
                if (!cfg.noprogress) progress_bar(i + 1, cfg.n_epochs, opt->last_energy_per_site(), 'S');
            }
            if (!cfg.noprogress) std::cout << std::endl;
            rbm.save(cfg.name);
        }
        if (cfg.evaluate && (rc == 0 || rc == 55)) {  // main.cpp:640-714
            rbmhost::setup ev(cfg);
            ev.use_evaluation_observables(true);
            if (cfg.sa_eval_samples != 0) ev.set_n_samples(cfg.sa_eval_samples);
            machine::gpu_rbm erbm(ev);
            erbm.set_params(rbm.params());
            sampler::gpu_metropolis_sampler es(erbm, ev);
            es.sample();
            std::cout.precision(16);
            const auto& gm = es.group_means();
            for (size_t i = 1; i < gm.size(); i++)
                std::cout << "Hex" << i - 1 << ": (" << gm[i].real() << "," << gm[i].imag() << ")" << std::endl;
            const double N = (double)rbm.n_visible;
            const auto& ah = es.energy();
            double stdv = ah.get_stddev() / N;
            std::cout << "Var: " << ah.get_variance()[0] / (N * N) << std::endl;
            std::cout << "Std: " << stdv << std::endl;
            std::cout << "Tau: " << ah.get_tau() << std::endl;
            std::cout << "Acc: " << es.get_acceptance_rate() << std::endl;
            std::cout << "E:   " << std::scientific << ah.get_result()[0].real() / N;
            std::cout << " \xC2\xB1 " << std::setprecision(1) << stdv << std::endl;
        }
    } catch (const std::exception& e) {
        std::cerr << "rbm: " << e.what() << std::endl;
        return 2;
    }
    return rc;
}
