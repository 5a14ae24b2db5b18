/*
 * rbmhost.h -- C ABI of the host layer above rbmgpu.h: builds lattice,
 * Hamiltonian and machine tables from the reference's .ini options
 * (src/tools/ini.cpp) the way src/main.cpp:125-239 does, and creates the
 * GPU context from them.  Used by language bindings (ctypes) and by the
 * `rbm` driver.
 */
#ifndef RBMHOST_H
#define RBMHOST_H
#include <stdint.h>

#include "rbmgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rbmhost_model rbmhost_model;

/* .ini text with the reference's keys; NULL on error (message in err) */
rbmhost_model* rbmhost_create(const char* ini_text, char* err, int32_t err_len);
void rbmhost_destroy(rbmhost_model* m);

/* [N, P, n_neural, n_vb, n_alpha, n_symm, pf_rows, n_terms, n_groups,
 *  n_bonds, n_chains, n_samples, n_warmup, step_size, n_prop_bonds,
 *  learn.hidden, cluster kmax] */
int rbmhost_info(const rbmhost_model* m, int64_t* out17);
int rbmhost_bonds(const rbmhost_model* m, int64_t* out /* 3 x n_bonds */);
int rbmhost_perms(const rbmhost_model* m, int64_t* out /* n_symm x N */);
/* cluster catalogue of the learned proposal (no counterpart in the
 * reference): N x kmax, row i = the first lattice loop through site i
 * (Kitaev hexagons, honeycomb.cpp:201-209; toric vertex stars,
 * toric_lattice.cpp:57-69) starting at i */
int rbmhost_clusters(const rbmhost_model* m, int64_t* out /* N x 8 */, int64_t* kmax);
/* sites (k each, concatenated) and 2^k x 2^k matrices of the terms */
int rbmhost_terms(const rbmhost_model* m, int64_t* k, int64_t* sites, double* mats, int64_t* groups);
double rbmhost_exact_energy(const rbmhost_model* m);
/* the Pfaffian index map bs_ (pfaffian.cpp:31-85): N x N column-major,
 * A_ij = sgn(bs_ij) fs_(|bs_ij| - 1, 2 s_i + s_j); rows = fs_ rows.
 * Returns -1 if the machine has no Pfaffian factor. */
int rbmhost_pf_map(const rbmhost_model* m, int64_t* out, int64_t* rows);

/* switch observables to the --evaluate set (main.cpp:640-680): helper
 * terms removed, plus one group per hexagon plaquette (Kitaev) */
int rbmhost_use_evaluation(rbmhost_model* m, int32_t hexagons);
/* sampler.eval_samples: n_samples = n, warmup rescaled (metropolis_sampler.cpp:289-293) */
int rbmhost_set_n_samples(rbmhost_model* m, int64_t n);

/* weights drawn like rbm_base::initialize_weights (std::mt19937(seed),
 * normal_distribution) followed by the Pfaffian weights */
int rbmhost_initial_weights(const rbmhost_model* m, uint64_t seed, double* params);

/* the massive-step guard of abstract_optimizer::optimize
 * (abstract_optimizer.cpp:157): 1 if max |dw_i| <= 30 and every entry is
 * finite, else 0 (the driver then exits with rc 55).  dw: n complex. */
int rbmhost_step_ok(const double* dw, int64_t n, double* max_abs);

/* update plugins of optimizer.plugin (plugin.cpp:30-77): kind "adam"
 * (p1..p3 = beta1, beta2, eps) or "momentum" (p1 = alpha, p2 = dialup).
 * apply turns the solver step dw (n complex, in place) into the update,
 * learning rate included, exactly as base_plugin::apply. */
typedef struct rbmhost_plugin rbmhost_plugin;
rbmhost_plugin* rbmhost_plugin_create(const char* kind, int64_t n, double p1, double p2, double p3);
int rbmhost_plugin_apply(rbmhost_plugin* pl, double* dw, int64_t n, double lr);
void rbmhost_plugin_destroy(rbmhost_plugin* pl);

/* GPU context for this model; precision / lanes / chains / samples of 0
 * keep the .ini values */
int rbmhost_gpu_create(const rbmhost_model* m, int32_t precision, int32_t lanes, rbmgpu_ctx** out);

#ifdef __cplusplus
}
#endif
#endif
