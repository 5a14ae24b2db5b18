/*
 * oracle.h -- C interface of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain restatement of the
 * reference's variational-Monte-Carlo path (dbucher97/rbm-nqs) used to check
 * the GPU library and to time the CPU baseline.  Only tests/, the smoke test
 * and bench.py's cpu_baseline leg load it; the product never links it.
 *
 * Conventions shared with the product:
 *   - a spin state is ceil(N/64) uint64 words, bit i set <=> sigma_i = -1
 *     (spin_state.cpp:44-50);
 *   - complex numbers are interleaved (re, im) doubles;
 *   - parameter vectors use the reference layout
 *     [v_bias | h_bias | W (column-major N x n_alpha) | Pfaffian fs_ (col-major)]
 *     (rbm_base.cpp:117-131, pfaffian.cpp:417-435);
 *   - the O matrix is P x S column-major, one column per sample
 *     (aggregator.cpp:274-280).
 * All functions return 0 on success and a negative value on error
 * (message from orc_error()).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

const char* orc_error(void);

/* Build lattice, Hamiltonian and machine from .ini text (reference keys). */
orc_model* orc_create(const char* ini_text);
void orc_destroy(orc_model* m);

/* sizes: [N, P, H(hidden units), n_symm, n_vb, n_alpha, n_pf, n_bonds,
 *         n_terms, n_words, n_neural] */
int orc_info(const orc_model* m, int64_t* out11);

/* lattice tables */
int orc_bonds(const orc_model* m, int64_t* out /* 3*n_bonds: a,b,type */);
int orc_nns(const orc_model* m, int64_t site, int64_t* out /* n_coord */);
int orc_hexagons(const orc_model* m, int64_t* out /* 6*n_uc_total */, int64_t* n);
int orc_plaqs(const orc_model* m, int64_t* out /* 5*n: 4 idx + type */, int64_t* n);
int orc_symmetry(const orc_model* m, int64_t* perms /* n_symm*N */);
/* Pfaffian index map bs_ (N x N column-major) and fs_ rows; -1 without a Pfaffian */
int orc_pf_map(const orc_model* m, int64_t* out, int64_t* rows);
int orc_up_down(const orc_model* m, int64_t uc, int64_t dir, int64_t step,
                int64_t* up, int64_t* down);
int orc_bravais_updown(int64_t n_uc, int64_t n_uc_b, int64_t h_shift, int64_t uc,
                       int64_t dir, int64_t step, int64_t* up, int64_t* down);
double orc_exact_energy(const orc_model* m);

/* Hamiltonian terms: for term t, sites and dense 2^k x 2^k matrix.
 * orc_term_info gives k for term t; orc_term returns sites + matrix. */
int orc_term_info(const orc_model* m, int64_t t, int64_t* k);
int orc_term(const orc_model* m, int64_t t, int64_t* sites, double* mat);

/* parameters (2*P doubles) */
int orc_set_params(orc_model* m, const double* params);
int orc_get_params(const orc_model* m, double* params);

/* per-state quantities (states: n * n_words uint64) */
int orc_theta(const orc_model* m, const uint64_t* state, double* theta /*2H*/);
int orc_logpsi(const orc_model* m, const uint64_t* states, int64_t n, double* out);
int orc_psi_ratio(const orc_model* m, const uint64_t* state, const int64_t* flips,
                  int64_t nf, double* out);
int orc_eloc(const orc_model* m, const uint64_t* states, int64_t n, double* out);
int orc_derivative(const orc_model* m, const uint64_t* states, int64_t n,
                   double* out /* P x n col-major complex */);
/* incremental theta updates along a flip sequence vs. from-scratch theta;
 * returns mean |dtheta| in *diff (test_derivative.cpp:53-76). */
int orc_update_check(const orc_model* m, const uint64_t* state,
                     const int64_t* flips, int64_t n_steps, int64_t per_step,
                     double* diff);

/* Pfaffian helpers */
int orc_pfaffian_dense(const double* A, int64_t n, double* mant /*2*/, int64_t* exp10);
int orc_pf_context(const orc_model* m, const uint64_t* state, double* inv /*2N^2*/,
                   double* mant, int64_t* exp10);
int orc_pf_update_check(const orc_model* m, const uint64_t* state,
                        const int64_t* flips, int64_t n_steps, int64_t per_step,
                        double* inv_diff, double* pf_diff, double* ratio_out);

/* Metropolis sampling with the reference's proposal semantics and
 * Philox-4x32-10 streams (key = seed, counter = (chain, epoch, step, slot)).
 * Outputs (may be NULL): states (S*n_words), eloc (2S), O (2*P*S),
 * chain acceptance (n_chains).  use_lut enables the reference's
 * state-keyed lncosh cache; threads = OpenMP threads (0 = default). */
int orc_sample(orc_model* m, uint64_t seed, uint32_t epoch, int threads, int use_lut,
               uint64_t* states, double* eloc, double* O, double* chain_acc);

/* Reference aggregator semantics on a fixed sample stream (unit weights).
 * e: 2S; O: 2*P*S (may be NULL).  out_e: [mean re, mean im, var, var_binned,
 * tau, stddev]; n_bins as in main.cpp:676-678 if 0.  out_o (6P): mean (2P),
 * var (P), <E O*> (2P), diag-unused (P). */
int orc_aggregate(const orc_model* m, const double* e, const double* O, int64_t S,
                  int64_t n_bins, double* out_e, double* out_o);

/* S.x (cg_solver.cpp:301-317) for an explicit O (P x S). */
int orc_apply_S(int64_t P, int64_t S, int64_t n_neural, const double* O,
                const double* mean, const double* var, double r1, double r2,
                double rd, const double* x, double* y);
/* cg1 (cg_solver.cpp:119-189) with warm start in x; returns itn, rs. */
int orc_cg(int64_t P, int64_t S, int64_t n_neural, const double* O,
           const double* mean, const double* var, double r1, double r2, double rd,
           const double* b, double* x, int64_t max_iter, double rtol,
           int64_t* itn, double* rs);

/* One full training epoch (sample + aggregate + SR + update) with the
 * reference optimiser semantics.  stats: [E/N, stddev/N, acc, tau, itn, rs,
 * max|dw|, ok, t_sample_s, t_opt_s]. */
int orc_epoch(orc_model* m, uint64_t seed, uint32_t epoch, int threads, int use_lut,
              double* stats);

/* exact expectation of E over |psi|^2 by enumerating 2^N states (N <= 24) */
int orc_exact_expectation(const orc_model* m, double* e_out /*2*/);

/* Philox-4x32-10 block (for stream tests) */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* learned proposal (no counterpart in the reference): cluster catalogue
 * cyc (N x kmax) and the FP64 proposal network, logits and batch gradient */
int orc_clusters(const orc_model* m, int64_t* cyc, int64_t* kmax);
int orc_mlp_logits(int64_t N, int64_t Hn, const double* W1, const double* b1, const double* W2,
                   const double* b2, const uint64_t* st, int64_t nw, int64_t n, double* z);
int orc_mlp_grad(int64_t N, int64_t Hn, const double* W1, const double* b1, const double* W2,
                 const double* b2, const uint64_t* st, int64_t nw, int64_t n, const int64_t* seed,
                 const double* w, double beta, double* gW1, double* gb1, double* gW2, double* gb2,
                 double* loss);

#ifdef __cplusplus
}
#endif
