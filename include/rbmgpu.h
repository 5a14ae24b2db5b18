/*
 * rbmgpu.h -- C ABI of the MI355X (gfx950) library for the variational
 * Monte Carlo epoch of rbm-nqs: RBM (+ Pfaffian) amplitudes, Metropolis
 * chains, local energies, O matrix, aggregation and the stochastic-
 * reconfiguration solve, all resident on one GPU.
 *
 * The reference (dbucher97/rbm-nqs) has no FFI: its hot path sits behind
 * C++ virtual interfaces.  Each entry point below names the reference
 * interface it replaces; the host adapters in rbm-nqs_amd/host implement
 * those interfaces on top of this ABI.
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles on the host side;
 *   - a spin state is n_words = ceil(N/64) uint64 words, bit i set <=>
 *     sigma_i = -1 (spin_state.cpp:44-50);
 *   - parameter vectors follow rbm_base.cpp:117-131 / pfaffian.cpp:417-435:
 *     [v_bias (n_vb) | h_bias (n_alpha) | W (N x n_alpha, column-major) |
 *      Pfaffian fs_ (pf_rows x 4, column-major)];
 *   - the O matrix is P x S column-major (one column per sample), like
 *     outer_aggregator_lazy (aggregator.cpp:270-280);
 *   - every function returns 0 on success and a negative code on failure;
 *     rbmgpu_last_error() gives the message.  No exceptions cross the ABI.
 *   - one context owns one HIP stream on the current device; a context is
 *     not thread-safe.
 */
#ifndef RBMGPU_H
#define RBMGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rbmgpu_ctx rbmgpu_ctx;

enum { RBMGPU_FP32 = 32, RBMGPU_FP64 = 64 };

typedef struct rbmgpu_config {
    /* ---- machine (rbm_base.cpp:38-52, rbm_symmetry.cpp:31-46) */
    int32_t n_spins;       /* N <= 128 (spin_state.hpp:20) */
    int32_t n_vb;          /* visible-bias size */
    int32_t n_alpha;       /* hidden rows n_alpha_ */
    int32_t n_symm;        /* symmetry permutations |G| (1 = basic RBM) */
    const int32_t* perm;   /* n_symm x N: perm[s*N+f] = indices()(f) */
    const int32_t* basis;  /* n_vb symmetry-basis sites (construct_symm_basis) */
    /* ---- Pfaffian factor (pfaffian.cpp:31-110); pf_rows = 0: none */
    int32_t pf_rows;
    const int32_t* pf_bs;  /* N x N column-major index map bs_ */
    int32_t pf_refresh;    /* sampler.pfaffian_refresh (abstract_sampler.cpp:103) */
    /* ---- observables: local operators (local_op.cpp:31-60) */
    int32_t n_terms;
    const int32_t* term_k;     /* sites per term, 1..8 */
    const int32_t* term_sites; /* concatenated site lists */
    const double* term_mat;    /* concatenated dense 2^k x 2^k row-major complex */
    const int32_t* term_group; /* observable index per term, non-decreasing;
                                  group 0 is the Hamiltonian */
    int32_t n_groups;
    /* ---- Metropolis sampler (metropolis_sampler.cpp:41-53) */
    int32_t n_chains;        /* 0: picked by the autotuner (no counterpart in the reference) */
    int64_t n_samples;
    int32_t n_warmup;
    int32_t step_size;
    double bond_flips;
    int32_t n_prop_bonds;
    const int32_t* prop_bonds; /* 2 x n_prop_bonds: (a, b) of non-z bonds */
    uint64_t seed;
    /* ---- execution */
    int32_t precision;       /* RBMGPU_FP32 or RBMGPU_FP64 */
    int32_t lanes_per_chain; /* 0: autotune; 8/16/32/64: lane-group sampler with that
                                many lanes per chain; -1: one chain per wavefront
                                (FP32 only) */
    int32_t keep_o;          /* 1: always materialise the O matrix.  0:
                                a basic RBM in FP32 uses the structured
                                S.x (O never formed, rbmgpu_get_samples
                                cannot return O); other machines
                                materialise O regardless */
    /* ---- learned proposal (no counterpart in the reference): cluster
     * flips drawn by a network z(s) = W2 relu(W1 s + b1) + b2 under
     * Metropolis-Hastings; FP32 machines without the Pfaffian factor */
    int32_t learn_hidden;     /* hidden width, multiple of 16, <= 512; 0: off */
    double learn_fraction;    /* probability of a learned move per slot */
    int32_t cluster_kmax;     /* cluster sizes 1 .. kmax (<= 8) */
    const int32_t* clusters;  /* N x kmax: the lattice loop through site i,
                                 starting at i */
    uint64_t learn_seed;      /* initial network weights */
    /* ---- SR epoch (rbmgpu_train_epoch): the CG's matrix-vector product
     * on the explicit S (1), matrix-free (0), or whichever the cost model
     * finds cheaper (-1, cg_solver.cpp:31-46) */
    int32_t sr_explicit_s;
} rbmgpu_config;

typedef struct rbmgpu_moments {
    double e_mean[2];   /* <E_loc> */
    double e_var;       /* unbiased variance (aggregator.cpp:145-152) */
    double e_var_binned;
    double e_tau;       /* binned tau (aggregator.cpp:120-128) */
    double e_std;       /* aggregator::get_stddev with binning */
    int64_t n_bins;
    double acceptance;  /* metropolis_sampler.cpp:89-92 */
} rbmgpu_moments;

/* metropolis_sampler / rbm_* / pfaffian constructors (main.cpp:153-239) */
int rbmgpu_create(const rbmgpu_config* cfg, rbmgpu_ctx** out);
void rbmgpu_destroy(rbmgpu_ctx* ctx);
const char* rbmgpu_last_error(const rbmgpu_ctx* ctx);
/* abstract_machine::get_n_params / get_n_neural_params (abstract_machine.hpp:86-101) */
int rbmgpu_n_params(const rbmgpu_ctx* ctx, int64_t* n_params, int64_t* n_neural);
/* chains the sampler runs (config n_chains = 0: the count the autotuner
 * picked at the first epoch; no counterpart in the reference) */
int rbmgpu_n_chains(const rbmgpu_ctx* ctx, int64_t* n);

/* execution layout (no counterpart in the reference): [lanes per chain,
 * units per lane, block size, flags: bit 0 S.x structured (no O matrix),
 * bit 1 the one-chain-per-wavefront sampler] */
int rbmgpu_launch_shape(const rbmgpu_ctx* ctx, int32_t* out4);

/* rbm_base::initialize_weights / load (rbm_base.cpp:58-113, 249-288) */
int rbmgpu_set_params(rbmgpu_ctx* ctx, const double* params);
/* rbm_base::save (rbm_base.cpp:219-247) */
int rbmgpu_get_params(const rbmgpu_ctx* ctx, double* params);
/* rbm_base::update_weights: w -= dw (rbm_base.cpp:117-143) */
int rbmgpu_update_weights(rbmgpu_ctx* ctx, const double* dw);

/* metropolis_sampler::sample + evaluate_and_aggregate for every kept sample
 * (metropolis_sampler.cpp:55-116, abstract_sampler.cpp:78-101): runs all
 * chains of one epoch on the GPU; states, local energies, tanh(theta) and
 * the O matrix (unless structured, see keep_o) stay on the device. */
int rbmgpu_sample_epoch(rbmgpu_ctx* ctx, uint32_t epoch);
/* operate on a host-given sample set instead (parity hook): evaluates
 * local observables, theta and O for these states as if they were sampled */
int rbmgpu_load_samples(rbmgpu_ctx* ctx, const uint64_t* states, int64_t n);

/* full_sampler::sample (full_sampler.cpp:43-185): the sample set becomes
 * all 2^N states (N <= 30) in the reference's order -- 2^parallel_bits
 * Gray-code runs over the upper bits, run b starting at the state b
 * (sampler.full.n_parallel_bits) -- each weighted by |psi|^2.  Moments, S.x
 * and the explicit S then use the weighted aggregator semantics
 * (aggregator.cpp:84-164, 270-305).  Needs the O matrix (keep_o). */
int rbmgpu_full_epoch(rbmgpu_ctx* ctx, int32_t parallel_bits);
/* log psi (2^N complex) and the weights |psi|^2 / max |psi|^2 (2^N) of the
 * last rbmgpu_full_epoch, in its state order; either pointer may be NULL */
int rbmgpu_full_weights(const rbmgpu_ctx* ctx, double* logpsi, double* w);

/* aggregator::finalize for the registered aggregators of SR
 * (abstract_optimizer.cpp:184-192, SR.cpp:74-79): E statistics, <O>,
 * Var(O), <E O*> and the means of the other observable groups.  Array
 * outputs may be NULL; sizes 2P, P, 2P, 2*n_groups. */
int rbmgpu_reduce(rbmgpu_ctx* ctx, rbmgpu_moments* out, double* o_mean,
                  double* o_var, double* eo_mean, double* group_means);

/* abstract_solver::solve via cg_solver::solve (cg_solver.cpp:260-330):
 * solves (S + reg) x = b with S from the last sampled / loaded set.
 * x holds the warm start on entry.  rtol = 0 -> 1e-16, max_iter = 0 ->
 * 2P (cg_solver.cpp:48-60).  method 0 = matrix-free CG on the device. */
int rbmgpu_sr_solve(rbmgpu_ctx* ctx, const double* b, double* x, double r1,
                    double r2, double rd, int64_t max_iter, double rtol,
                    int64_t* itn, double* rs);
/* One whole SR epoch on the device (main.cpp:596-610: sampler->sample()
 * then optimizer->optimize()): sampling, local energies, moments, F =
 * <E O*> - <E><O>* (SR.cpp:102), the CG solve warm-started from the
 * previous epoch's lr-scaled step (cg_solver.cpp:119-189, SR.cpp:57), the
 * massive-step guard (abstract_optimizer.cpp:157; non-finite steps are
 * refused too) and w -= lr x (rbm_base.cpp:117-143).  The decayed
 * lr / reg values of this epoch (abstract_optimizer.cpp:39-48) are the
 * caller's.  With graph = 1 the body is captured once in a HIP graph and
 * replayed (no counterpart in the reference); the parameters then live on
 * the device (rbmgpu_get_params reads them from there). */
typedef struct rbmgpu_sr_params {
    double lr, reg1, reg2, reg1_delta;
    int64_t max_iter;  /* 0: 2P (cg_solver.cpp:48-56) */
    double rtol;       /* 0: 1e-16 */
    int32_t graph;     /* 1: replay a captured HIP graph, 0: launch eagerly */
} rbmgpu_sr_params;
typedef struct rbmgpu_epoch_stats {
    double e_mean[2], e_var, e_var_binned, e_tau, e_std;
    int64_t n_bins;
    double acceptance;
    int64_t cg_itn;
    double cg_rs;
    double max_dw;        /* max |x| of the solve (NaN if not finite) */
    int32_t updated;      /* 0: the guard refused the step (driver rc 55) */
    int32_t graph_replayed;
    double ms;            /* host wall time of the call */
    int32_t explicit_s;   /* 1: the CG ran on the explicit S (cheaper than
                             matrix-free for this sample set, cg_solver.cpp:31-46) */
    int32_t pad_;
} rbmgpu_epoch_stats;
int rbmgpu_train_epoch(rbmgpu_ctx* ctx, uint32_t epoch, const rbmgpu_sr_params* p, rbmgpu_epoch_stats* out);

/* the SR solver family (SR.cpp:51-72): RBMGPU_SOLVE_CG is the matrix-free
 * CG of rbmgpu_sr_solve (the reference's "cg" and "minres",
 * cg_solver.cpp:31-37); RBMGPU_SOLVE_CG_SINGLE builds S explicitly
 * (cg_solver.cpp:274-299) and runs the same CG on it; RBMGPU_SOLVE_DIRECT
 * builds S with the direct solver's multiplicative regularisation and
 * solves it exactly (direct_solver.cpp:31-59; itn = 0, rs = |S x - b|^2).
 * The explicit forms need the O matrix (not the structured S.x). */
enum { RBMGPU_SOLVE_CG = 0, RBMGPU_SOLVE_CG_SINGLE = 1, RBMGPU_SOLVE_DIRECT = 2 };
int rbmgpu_sr_solve_ex(rbmgpu_ctx* ctx, int32_t method, const double* b, double* x, double r1, double r2,
                       double rd, int64_t max_iter, double rtol, int64_t* itn, double* rs);
/* the explicit S of the current sample set with the CG regularisation of
 * cg_solver.cpp:276-281 (P x P column-major complex); parity hook */
int rbmgpu_build_S(rbmgpu_ctx* ctx, double r1, double r2, double rd, double* S);

/* one S.x product (the Aprod lambda of cg_solver.cpp:301-317) */
int rbmgpu_apply_S(rbmgpu_ctx* ctx, const double* x, double* y, double r1,
                   double r2, double rd);

/* Allocation-free variant of rbmgpu_apply_S.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  x and y are DEVICE
 * pointers (P complex doubles each); the call is asynchronous on the
 * context's stream (rbmgpu_synchronize waits for it).  The S.x block
 * partials go to a static per-device scratch of 8 MiB (at P = 2144 in FP32
 * with the structured S.x about 4.5 MB); a sample set whose partials need
 * more is rejected with an error.  The moments of the current sample set
 * must exist (rbmgpu_reduce first).  Every call on one GPU shares that
 * scratch, so calls that overlap on one GPU (several contexts or streams)
 * race. */
int rbmgpu_apply_S_noalloc(rbmgpu_ctx* ctx, const double* x, double* y, double r1, double r2, double rd);

/* parity hooks on host-given states */
int rbmgpu_eval_logpsi(rbmgpu_ctx* ctx, const uint64_t* states, int64_t n,
                       double* logpsi);
/* abstract_machine::psi_over_psi (abstract_machine.hpp:197-214) */
int rbmgpu_psi_ratio(rbmgpu_ctx* ctx, const uint64_t* state, const int32_t* flips,
                     int32_t n_flips, double* ratio);
/* copies of the current sample set: states (S x n_words), local
 * observables (S x n_groups complex), O (P x S complex; not in
 * structured mode),
 * per-chain acceptance (n_chains).  Any pointer may be NULL. */
int rbmgpu_get_samples(const rbmgpu_ctx* ctx, uint64_t* states, double* obs,
                       double* O, double* chain_acc);
int rbmgpu_n_samples(const rbmgpu_ctx* ctx, int64_t* n);

/* device time of the last launch of each kernel family, from HIP events
 * on the context's stream (no counterpart in the reference): [sample,
 * obs/O assembly, reduce, one CG iteration (S.x and the CG step; mean over
 * the last rbmgpu_sr_solve, launches queued past convergence included) or
 * the last rbmgpu_apply_S, unused (0), S.x main kernel alone (k_struct or
 * k_sx_partial, mean of the last rbmgpu_time_sx)] in ms. */
int rbmgpu_kernel_times(const rbmgpu_ctx* ctx, double* ms6);
int rbmgpu_synchronize(rbmgpu_ctx* ctx);
/* ---- learned proposal (no counterpart in the reference) */
/* one training step of the proposal network on the records of the last
 * rbmgpu_sample_epoch (the last learned proposal of each chain before each
 * kept sample): REINFORCE with reward a * k (acceptance probability times
 * cluster size, centred on the batch mean) and entropy weight beta, Adam
 * with learning rate lr on at most batch records.  out4: [loss, mean
 * reward, learned moves proposed, accepted] of the epoch. */
int rbmgpu_learn_train(rbmgpu_ctx* ctx, double lr, double beta, int64_t batch, double* out4);
/* batch objective and gradient at the current weights for host-given
 * records (parity hook): grads = [W1 (Hn x N) | b1 (Hn) | W2 (N x Hn) |
 * b2 (N)] */
int rbmgpu_learn_grad(rbmgpu_ctx* ctx, const uint64_t* states, const int32_t* seeds, const double* w,
                      int64_t n, double beta, double* grads, double* loss);
/* FP32 master weights in the same layout; set = 1 uploads them */
int rbmgpu_learn_weights(rbmgpu_ctx* ctx, int32_t set, double* params);
/* logits of host-given states through the sampler's own forward (per-wavefront FP32
   for networks of at most 64 sites and hidden units, else bf16 MFMA) */
int rbmgpu_learn_logits(rbmgpu_ctx* ctx, const uint64_t* states, int64_t n, double* z);
/* learn_fraction for the following epochs (0: plain Metropolis) */
int rbmgpu_learn_set_fraction(rbmgpu_ctx* ctx, double fraction);

/* mean device time of the main S.x kernel (k_struct or k_sx_partial) over
 * reps back-to-back launches on the current sample set, bracketed by two
 * HIP events (no counterpart in the reference; measurement hook) */
int rbmgpu_time_sx(rbmgpu_ctx* ctx, int32_t reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
