// capi.cpp -- implementation of include/rbmhost.h.
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/rbmhost.h"
#include "optim.hpp"
#include "setup.hpp"

struct rbmhost_model {
    std::unique_ptr<rbmhost::setup> s;
};

extern "C" {

rbmhost_model* rbmhost_create(const char* ini_text, char* err, int32_t err_len) {
    try {
        ini::config cfg;
        ini::parse_ini_text(ini_text ? ini_text : "", cfg);
        ini::finalize(cfg, !cfg.name.empty());
        auto* m = new rbmhost_model();
        m->s.reset(new rbmhost::setup(cfg));
        return m;
    } catch (const std::exception& e) {
        if (err && err_len > 0) {
            std::strncpy(err, e.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return nullptr;
    }
}

void rbmhost_destroy(rbmhost_model* m) { delete m; }

int rbmhost_info(const rbmhost_model* m, int64_t* o) {
    if (!m || !o) return -1;
    const auto& s = *m->s;
    const auto& L = s.layout();
    const auto& cfg = s.cfg();
    o[0] = L.n_spins;
    o[1] = L.n_params();
    o[2] = L.n_neural();
    o[3] = L.n_vb;
    o[4] = L.n_alpha;
    o[5] = L.n_symm;
    o[6] = L.pf_rows;
    o[7] = (int64_t)s.terms().k.size();
    o[8] = s.terms().n_groups;
    o[9] = (int64_t)m->s->model().get_lattice().get_bonds().size();
    o[10] = (int64_t)cfg.sa_metropolis_n_chains;
    o[11] = (int64_t)s.n_samples();
    o[12] = (int64_t)s.warmup();
    o[13] = (int64_t)cfg.sa_metropolis_n_steps_per_sample;
    o[14] = (int64_t)s.prop_bonds().size() / 2;
    o[15] = cfg.learn_hidden;
    o[16] = s.cluster_kmax();
    return 0;
}

int rbmhost_bonds(const rbmhost_model* m, int64_t* out) {
    if (!m || !out) return -1;
    size_t i = 0;
    for (const auto& b : m->s->model().get_lattice().get_bonds()) {
        out[i++] = (int64_t)b.a;
        out[i++] = (int64_t)b.b;
        out[i++] = (int64_t)b.type;
    }
    return 0;
}

int rbmhost_perms(const rbmhost_model* m, int64_t* out) {
    if (!m || !out) return -1;
    const auto& p = m->s->layout().perm;
    for (size_t i = 0; i < p.size(); i++) out[i] = p[i];
    return 0;
}

int rbmhost_clusters(const rbmhost_model* m, int64_t* out, int64_t* kmax) {
    if (!m || !out || !kmax) return -1;
    const auto& c = m->s->clusters();
    for (size_t i = 0; i < c.size(); i++) out[i] = c[i];
    *kmax = m->s->cluster_kmax();
    return 0;
}

int rbmhost_terms(const rbmhost_model* m, int64_t* k, int64_t* sites, double* mats, int64_t* groups) {
    if (!m) return -1;
    const auto& t = m->s->terms();
    for (size_t i = 0; i < t.k.size(); i++) {
        if (k) k[i] = t.k[i];
        if (groups) groups[i] = t.group[i];
    }
    if (sites)
        for (size_t i = 0; i < t.sites.size(); i++) sites[i] = t.sites[i];
    if (mats) std::memcpy(mats, t.mat.data(), t.mat.size() * sizeof(double));
    return 0;
}

double rbmhost_exact_energy(const rbmhost_model* m) {
    if (!m) return 0;
    if (m->s->cfg().model == "kitaev") return static_cast<model::kitaev&>(m->s->model()).exact_energy();
    return m->s->cfg().J.at(0);
}

int rbmhost_use_evaluation(rbmhost_model* m, int32_t hexagons) {
    if (!m) return -1;
    m->s->use_evaluation_observables(hexagons != 0);
    return 0;
}

int rbmhost_set_n_samples(rbmhost_model* m, int64_t n) {
    if (!m || n < 1) return -1;
    m->s->set_n_samples((size_t)n);
    return 0;
}

int rbmhost_initial_weights(const rbmhost_model* m, uint64_t seed, double* params) {
    if (!m || !params) return -1;
    std::mt19937 rng((std::mt19937::result_type)seed);
    auto p = m->s->initial_weights(rng);
    std::memcpy(params, p.data(), p.size() * sizeof(std::complex<double>));
    return 0;
}

int rbmhost_gpu_create(const rbmhost_model* m, int32_t precision, int32_t lanes, rbmgpu_ctx** out) {
    if (!m || !out) return -1;
    rbmgpu_config c = m->s->gpu_config();
    if (precision) c.precision = precision;
    if (lanes) c.lanes_per_chain = lanes;
    return rbmgpu_create(&c, out);
}

int rbmhost_pf_map(const rbmhost_model* m, int64_t* out, int64_t* rows) {
    if (!m || !out || !rows) return -1;
    const auto& L = m->s->layout();
    if (!L.pf_rows) return -1;
    for (size_t i = 0; i < L.pf_bs.size(); i++) out[i] = L.pf_bs[i];
    *rows = L.pf_rows;
    return 0;
}

rbmhost_plugin* rbmhost_plugin_create(const char* kind, int64_t n, double p1, double p2, double p3) {
    if (!kind || n < 0) return nullptr;
    const std::string k(kind);
    if (k == "adam") return reinterpret_cast<rbmhost_plugin*>(new optim::adam((size_t)n, p1, p2, p3));
    if (k == "momentum") return reinterpret_cast<rbmhost_plugin*>(new optim::momentum((size_t)n, p1, p2));
    return nullptr;
}

int rbmhost_plugin_apply(rbmhost_plugin* pl, double* dw, int64_t n, double lr) {
    if (!pl || !dw || n < 0) return -1;
    reinterpret_cast<optim::plugin*>(pl)->apply(reinterpret_cast<std::complex<double>*>(dw), (size_t)n, lr);
    return 0;
}

void rbmhost_plugin_destroy(rbmhost_plugin* pl) { delete reinterpret_cast<optim::plugin*>(pl); }

int rbmhost_step_ok(const double* dw, int64_t n, double* max_abs) {
    if (!dw || n < 0) return -1;
    return optim::step_ok((const std::complex<double>*)dw, (size_t)n, max_abs) ? 1 : 0;
}

}  // extern "C"
