// setup.cpp -- see setup.hpp.
#include "setup.hpp"

#include <algorithm>
#include <stdexcept>

namespace rbmhost {

void term_table::add(const model::local_op& op, int g) {
    k.push_back((int32_t)op.acts_on.size());
    for (size_t s : op.acts_on) sites.push_back((int32_t)s);
    group.push_back(g);
    for (const auto& v : op.mat) {
        mat.push_back(v.real());
        mat.push_back(v.imag());
    }
    n_groups = std::max(n_groups, g + 1);
}

// Index map of the Pfaffian's skew matrix under the lattice symmetry
// (pfaffian.cpp:31-85): entry (i,j) of A(sigma) is sgn(bs_ij) *
// fs_(|bs_ij| - 1, 2 s_i + s_j); equivalent entries share a row of fs_.
static void pfaffian_index_map(const lattice::bravais& lat, const std::vector<double>& symm,
                               std::vector<int32_t>& bs, int& rows) {
    const size_t n = lat.n_total;
    const auto perms = lat.construct_symmetry(symm);
    const auto basis = lat.construct_symm_basis(symm);
    bs.assign(n * n, 0);
    auto at = [&](size_t i, size_t j) -> int32_t& { return bs[i + j * n]; };
    int offset = 0;
    for (size_t bi : basis) {
        for (const auto& p : perms) {
            const int mirrored = p[0] > p[1] ? -1 : 1;  // orientation-reversing symmetry
            for (size_t j = 0; j < n; j++) {
                const size_t a = p[bi], b = p[j];
                at(b, a) = (a == b) ? 0 : mirrored * (int)(j + offset);
            }
        }
        offset += (int)n;
    }
    // antisymmetry: keep the entry of smaller magnitude on both sides
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < i; j++) {
            if (std::abs(at(i, j)) <= std::abs(at(j, i))) at(j, i) = -at(i, j);
            else at(i, j) = -at(j, i);
        }
    // relabel the used magnitudes densely 1..rows, preserving order
    auto maxv = [&]() { return *std::max_element(bs.begin(), bs.end()); };
    for (int v = 1; v < maxv(); v++) {
        while (std::count(bs.begin(), bs.end(), v) == 0) {
            for (auto& x : bs) {
                if (x > v) x--;
                else if (x < -v) x++;
            }
        }
    }
    rows = maxv();
}

setup::setup(const ini::config& cfg) : cfg_(cfg) {
    if (cfg.model == "kitaev") {
        if (cfg.lattice_type == "hex") throw std::runtime_error("honeycomb_hex lattices are not supported");
        auto* m = new model::kitaev(cfg.n_cells, cfg.J, cfg.n_cells_b, cfg.symmetry, cfg.h);
        model_.reset(m);
        if (cfg.helper_strength != 0.) m->add_helper_hamiltonian(cfg.helper_strength);
    } else if (cfg.model == "toric") {
        model_.reset(new model::toric(cfg.n_cells, cfg.J.at(0)));
    } else {
        throw std::runtime_error("model '" + cfg.model + "' is not supported by the GPU backend");
    }
    auto& lat = model_->get_lattice();
    layout_.n_spins = (int)lat.n_total;
    if (lat.n_total > 128) throw std::runtime_error("n must be smaller than MAX_SPIN_SITES!");
    std::vector<lattice::perm_t> perms;
    std::vector<size_t> basis;
    if (cfg.rbm == "basic") {  // rbm_base(alpha, lattice): n_vb = N
        layout_.n_vb = layout_.n_spins;
        layout_.n_symm = 1;
        perms = lat.construct_symmetry({0});
        basis = lat.construct_symm_basis({0});
    } else if (cfg.rbm == "symmetry") {  // rbm_symmetry.cpp:31-46
        const auto& sy = lat.default_symmetry();
        layout_.n_symm = (int)lat.symmetry_size(sy);
        layout_.n_vb = layout_.n_spins / layout_.n_symm;
        perms = lat.construct_symmetry(sy);
        basis = lat.construct_symm_basis(sy);
        if ((int)perms.size() != layout_.n_symm) throw std::runtime_error("symmetry size mismatch");
    } else {
        throw std::runtime_error("rbm.type '" + cfg.rbm + "' is not supported by the GPU backend");
    }
    layout_.n_alpha = (int)cfg.alpha * layout_.n_vb;
    for (auto& p : perms)
        for (size_t v : p) layout_.perm.push_back((int32_t)v);
    for (size_t b : basis) layout_.basis.push_back((int32_t)b);
    if (cfg.rbm_pfaffian) pfaffian_index_map(lat, cfg.rbm_pfaffian_symmetry, layout_.pf_bs, layout_.pf_rows);
    for (const auto& b : lat.get_bonds())
        if (b.type != 2) {  // metropolis_sampler.cpp:157-166
            prop_bonds_.push_back((int32_t)b.a);
            prop_bonds_.push_back((int32_t)b.b);
        }
    for (const auto& op : model_->get_hamiltonian()) terms_.add(op, 0);
    {  // cluster catalogue (learned proposal)
        std::vector<std::vector<size_t>> loops;
        if (cfg.model == "kitaev") {
            loops = static_cast<model::kitaev*>(model_.get())->honeycomb_lattice().get_hexagons();
        } else {
            for (const auto& p : static_cast<lattice::toric_lattice&>(lat).construct_plaqs())
                if (p.type == 1) loops.push_back({p.idxs.begin(), p.idxs.end()});
        }
        cluster_kmax_ = loops.empty() ? 0 : (int)loops[0].size();
        for (int i = 0; i < layout_.n_spins; i++) {
            bool found = false;
            for (const auto& l : loops) {
                auto it = std::find(l.begin(), l.end(), (size_t)i);
                if (it == l.end()) continue;
                size_t p0 = (size_t)(it - l.begin());
                for (int t = 0; t < cluster_kmax_; t++) clusters_.push_back((int32_t)l[(p0 + t) % l.size()]);
                found = true;
                break;
            }
            if (!found) throw std::runtime_error("site on no lattice loop");
        }
    }
    n_samples_ = cfg.sa_n_samples;
    warmup_ = cfg.sa_metropolis_n_warmup_steps;
}

void setup::use_evaluation_observables(bool hexagons) {
    model_->remove_helper_hamiltonian();
    terms_ = term_table();
    for (const auto& op : model_->get_hamiltonian()) terms_.add(op, 0);
    if (hexagons && cfg_.model == "kitaev") {
        auto* k = static_cast<model::kitaev*>(model_.get());
        const auto plaq = model::plaquette_op();
        int g = 1;
        for (const auto& hx : k->honeycomb_lattice().get_hexagons()) terms_.add({hx, plaq.n, plaq.a}, g++);
    }
}

void setup::set_n_samples(size_t n) {
    double f = n / (double)n_samples_;
    warmup_ = (size_t)(warmup_ * f);
    n_samples_ = n;
}

rbmgpu_config setup::gpu_config() const {
    rbmgpu_config c{};
    c.n_spins = layout_.n_spins;
    c.n_vb = layout_.n_vb;
    c.n_alpha = layout_.n_alpha;
    c.n_symm = layout_.n_symm;
    c.perm = layout_.perm.data();
    c.basis = layout_.basis.data();
    c.pf_rows = layout_.pf_rows;
    c.pf_bs = layout_.pf_bs.empty() ? nullptr : layout_.pf_bs.data();
    c.pf_refresh = cfg_.sa_pfaffian_refresh;
    c.n_terms = (int32_t)terms_.k.size();
    c.term_k = terms_.k.data();
    c.term_sites = terms_.sites.data();
    c.term_mat = terms_.mat.data();
    c.term_group = terms_.group.data();
    c.n_groups = terms_.n_groups;
    c.n_chains = cfg_.gpu_auto_chains && cfg_.sa_type != "full" ? 0 : (int32_t)cfg_.sa_metropolis_n_chains;
    c.n_samples = (int64_t)n_samples_;
    c.n_warmup = (int32_t)warmup_;
    c.step_size = (int32_t)cfg_.sa_metropolis_n_steps_per_sample;
    c.bond_flips = cfg_.sa_metropolis_bond_flips;
    c.n_prop_bonds = (int32_t)(prop_bonds_.size() / 2);
    c.prop_bonds = prop_bonds_.data();
    c.seed = cfg_.seed;
    c.precision = cfg_.gpu_precision;
    c.lanes_per_chain = cfg_.gpu_lanes;
    // the explicit-S solvers (cg-single, direct and possibly cg-dynamic) read O
    const auto& m = cfg_.opt_sr_method;
    const bool explicit_S = cfg_.opt_type == "SR" && (m == "cg-single" || m == "direct" || m == "cg-dynamic");
    // exact enumeration weights the samples, which the structured S.x does not do
    c.keep_o = (cfg_.gpu_keep_o || explicit_S || cfg_.sa_type == "full") ? 1 : 0;
    c.sr_explicit_s = cfg_.gpu_explicit_s;
    c.learn_hidden = cfg_.learn_hidden;
    c.learn_fraction = cfg_.learn_fraction;
    c.cluster_kmax = cluster_kmax_;
    c.clusters = clusters_.data();
    c.learn_seed = cfg_.seed;
    return c;
}

std::vector<std::complex<double>> setup::initial_weights(std::mt19937& rng) const {
    double sd = cfg_.rbm_weights, si = cfg_.rbm_weights_imag < 0 ? sd : cfg_.rbm_weights_imag;
    std::normal_distribution<double> real_dist{0, sd}, imag_dist{0, si};
    const int N = layout_.n_spins, nvb = layout_.n_vb, na = layout_.n_alpha;
    std::vector<std::complex<double>> p(layout_.n_params(), 0.);
    if (nvb == N && cfg_.model == "kitaev" && !cfg_.rbm_weights_init_type.empty()) {
        // honeycomb::initialize_vb (honeycomb.cpp:179-191, rbm_base.cpp:72-75):
        // no draws for v_bias; "stripy" puts i pi/2 on every site of the unit
        // cells u with ((u / n_uc_b) % 2 + u % n_uc_b) % 2 == 0
        if (cfg_.rbm_weights_init_type == "stripy") {
            const auto& lat = model_->get_lattice();
            for (int i = 0; i < N; i++) {
                const size_t u = lat.uc_idx(i);
                if (((u / lat.n_uc_b) % 2 + u % lat.n_uc_b) % 2 == 0) p[i] = std::complex<double>(0, 1.57079632679);
            }
        }
    } else {
        for (int i = 0; i < nvb; i++) p[i] = std::complex<double>(real_dist(rng), imag_dist(rng));
    }
    for (int i = 0; i < na; i++) {
        p[nvb + i] = std::complex<double>(real_dist(rng), imag_dist(rng));
        for (int j = 0; j < N; j++)
            p[nvb + na + j + (int64_t)i * N] = std::complex<double>(real_dist(rng), imag_dist(rng));
    }
    if (layout_.pf_rows) {
        std::normal_distribution<double> dist{0, cfg_.rbm_pfaffian_weights};
        for (int64_t i = 0; i < 4 * layout_.pf_rows; i++)
            p[layout_.n_neural() + i] = std::complex<double>(dist(rng), dist(rng));
    }
    return p;
}

}  // namespace rbmhost
