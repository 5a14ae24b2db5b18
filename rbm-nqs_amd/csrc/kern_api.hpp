// kern_api.hpp -- argument blocks of the gfx950 kernels and the host-side
// launchers (one instantiation per precision in kern_f32.hip / kern_f64.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.hpp"
#include "kern_pf.hpp"

namespace rbmgpu {

struct Dims {
    int N, H, n_alpha, n_symm, n_vb, nw;
    int w_lds;  // 1: the W table is staged in LDS too (else read from memory)
    int t_lds;  // 1: the local-operator tables are staged in LDS too
    int npb;    // proposal bonds
    int n_terms, n_rowptr, n_ent;  // sizes of the local-operator tables
    int pf_rows;  // Pfaffian fs_ rows (0: no Pfaffian factor)
    int waves;    // wavefronts per block (per-wave Pfaffian scratch)
};

// weight tables, all complex R; index k*n_alpha + alpha for the N x n_alpha
// tables (k = permuted site)
template <typename R>
struct Tabs {
    const cplx<R>* W;
    const cplx<R>* C;   // cosh(2W)
    const cplx<R>* S;   // sinh(2W)
    const cplx<R>* hb;  // hidden bias [n_alpha]
    const cplx<R>* rb;  // ratio bias per site: v_{f mod n_vb} (rbm_base.cpp:310)
    const cplx<R>* lb;  // log-psi bias per site: sum_i v_i #{s: perm_s(basis_i) = f}
    const uint8_t* perm;
    const cplx<R>* fs;  // Pfaffian weights, pf_rows x 4 (fs[r*4 + c])
    const int16_t* bs;  // Pfaffian index map, N x N (bs[i*N + j])
};

// local operators as sparse rows: entries [rowptr[rowbase[t] + loc],
// rowptr[rowbase[t] + loc + 1]) hold (xor mask, value) of row loc of term t
struct Terms {
    int n_terms, n_groups;
    const int32_t* k;
    const uint8_t* sites;  // n_terms x 8
    const int32_t* group;
    const int32_t* rowbase;
    const int32_t* rowptr;
    const uint8_t* mask;
    const double* val;     // complex
};

template <typename R>
struct SampArgs {
    Dims d;
    Tabs<R> t;
    Terms T;
    int n_chains;
    int64_t n_samples;
    int warmup, step, npb;
    uint64_t thr;  // bond flip if philox word < thr (thr = bond_flips * 2^32)
    uint32_t epoch, k0, k1;
    const uint32_t* epoch_dev;  // non-null: the epoch is read from device memory (graph replays)
    int pf_refresh;  // fresh Pfaffian inverse every pf_refresh accepted moves (0: never)
    int obs_later;   // 1: kept samples store state and tanh only, k_eloc adds the observables
    const int32_t* pb;
    uint64_t* states;
    double* obs;
    cplx<R>* Tout;
    cplx<R>* Opf;    // S x 4 pf_rows: d ln Pf / d fs per kept sample
    uint32_t* acc;
    uint32_t* prop;
    // k_sample_wave, symmetric RBM: image pairs of the proposal bonds
    const uint8_t* pmap;   // n_symm x npb: ordered image pair of bond b under symmetry s
    const int32_t* pairs;  // npair x 2: (lead, other) site of each image pair
    int npair;             // 0: bond moves composed from the single-site tables
    const int16_t* tbond;  // per term: the proposal bond of its two sites, -1
};

// local observables of stored samples (k_eloc): FP32 basic RBM, terms of
// at most 4 sites, N x H tables (cosh 2W, sinh 2W at [f * H + j])
constexpr int kElocThreads = 512;  // one block's LDS tables serve 8 wavefronts
struct ElocArgs {
    int64_t n;  // samples
    int N, H, nw;
    Terms T;
    const cplx<float>* C;
    const cplx<float>* S;
    const cplx<float>* rb;
    const uint64_t* states;  // n x nw
    const cplx<float>* tanh;  // n x H
    double* obs;              // n x 2 n_groups
    int n_ratio;              // terms with an off-diagonal entry
    const int32_t* ratio_term;  // [n_ratio] their indices, in order
    const int32_t* term_slot;   // [n_terms] index into ratio_term, or -1 (diagonal only)
};
size_t eloc_lds_bytes(int N, int H);
int eloc_blocks_per_cu(int N, int H);
hipError_t launch_eloc(const ElocArgs& a, int blocks, hipStream_t st);

template <typename R>
struct EvalArgs {
    Dims d;
    Tabs<R> t;
    Terms T;
    int64_t n;
    const uint64_t* states;
    double* logpsi;  // optional
    double* obs;     // optional
    cplx<R>* Tout;   // optional
    double* ratio;   // optional, with flips
    cplx<R>* Opf;    // optional (Pfaffian): n x 4 pf_rows
    const int32_t* pb;
    int32_t flips[8];
    int n_flips;
};

template <typename R>
struct OArgs {
    Dims d;
    int64_t S;
    int P, P_rbm, ld;
    const uint64_t* states;
    const cplx<R>* T;
    const int32_t* vb_site;  // n_symm x n_vb: perm_s(basis_i)
    const int32_t* invperm;  // n_symm x N
    const cplx<R>* Opf;      // S x (P - P_rbm) Pfaffian block, or null
    cplx<R>* O;
};

template <typename R>
struct MomArgs {
    int64_t S, chunk;
    int P, n_groups, ld;
    const cplx<R>* O;
    const double* obs;
    double* part;      // blocks x P x kOPartRow (k_o_partial)
    const double* w;   // per-sample weights (exact enumeration); nullptr: 1
};
constexpr int kOPartRow = 9;

struct CgStatus {
    double rsold, rs;
    int64_t it, itn, max_iter;
    double rtol;
    int done;
};

template <typename R>
struct SxArgs {
    int64_t S, nblocks;
    int P, n_neural, ld, nchunks;
    const cplx<R>* O;
    const double* x;     // complex P
    const double* mean;  // complex P
    const double* var;   // P
    const double* maxvar;
    double r1, r2, rd;
    const double* reg_dev;  // non-null: r1, r2, rd read from device memory (graph replays)
    const double* wsx;      // non-null: per-sample weight S w_k / sum w of an enumerated set
    cplx<R>* part;       // nblocks x ld
    double* part2;       // nchunks x P complex
    double* y;           // complex P
    double* pdot;        // per column block: p^H y partial (2 doubles), optional
    const CgStatus* status;
};

// learned proposal (no counterpart in the reference); see kern_learn.hpp
struct LearnArgs {
    SampArgs<float> s;
    int Hn, Np, kmax;        // hidden width, N padded to 16, loop length
    uint32_t lam_thr;        // learned move if the Philox word < lam_thr
    const uint16_t* W1;      // bf16, Hn x Np (row h: weights of hidden unit h)
    const uint16_t* W2;      // bf16, Np x Hn
    const float* b1;         // Hn
    const float* b2;         // Np
    const uint8_t* cyc;      // N x kmax: cluster (i, k) = cyc[i][0..k-1]
    const uint8_t* seedc;    // N x kmax: seeds generating the set of (i, k)
    const uint8_t* seedl;    // N x kmax x 6
    uint64_t* rec_st;        // per kept sample: state before the last learned proposal
    int16_t* rec_seed;       // its seed (-1: none)
    uint8_t* rec_k;          // its size
    float* rec_a;            // its acceptance probability
    uint32_t* lacc;          // per chain: accepted / proposed learned moves
    uint32_t* lprop;
};
// small networks (hidden width and sites at most 64) are evaluated by each
// chain's wavefront on its own (FP32 VALU dot products of the bf16 weights,
// no block synchronisation); larger ones by the block's 8 chains together
// on MFMA
__host__ __device__ inline bool mlp_wave_fwd(int Hn, int N) { return Hn <= 64 && N <= 64; }
// LDS bytes of the proposal network region of k_sample_mlp (after the
// sampler tables)
__host__ __device__ inline size_t mlp_lds_bytes(int Hn, int Np, int N, int kmax) {
    size_t b = 2 * (size_t)Hn * Np * 2 + ((size_t)Hn + Np) * 4;  // W1, W2 bf16, b1, b2
    b += 16 * (size_t)Np * 2 + 16 * (size_t)Hn * 2 + 2 * 16 * (size_t)Np * 4;  // Sg, Ha, Zt, Zc
    b += (size_t)N * kmax * 8;  // cyc, seedc, seedl
    b = (b + 15) & ~(size_t)15;
    if (mlp_wave_fwd(Hn, N)) b += 2 * (size_t)Hn * Np * 4 + 16 * (size_t)Hn * 4;  // W1^T, W2^T FP32, hidden rows
    return (b + 15) & ~(size_t)15;
}
// chains (wavefronts) per k_sample_mlp block: at most kMlpChains for the
// unconstrained kernel (two waves per SIMD at its register count), up to
// kMlpChainsMax (the 16 rows of the MFMA tiles) for the 128-VGPR variant;
// MFMA rows beyond the block's chains stay zero
constexpr int kMlpChains = 8;
constexpr int kMlpChainsMax = 16;
hipError_t launch_sample_mlp(const LearnArgs& a, int K, hipStream_t st);
hipError_t launch_mlp_logits(const LearnArgs& a, const uint64_t* states, int64_t n, float* z, hipStream_t st);

// training of the proposal network on a batch of records (FP32)
struct TrainArgs {
    int N, Np, Hn, nw;
    int64_t B;
    const uint64_t* st;      // B x nw states
    const int16_t* seed;     // B
    const float* w;          // B rewards (already centred)
    float beta, inv_b;       // entropy weight, 1 / B
    const float* W1;         // fp32 masters, Hn x N (row-major)
    const float* b1;
    const float* W2;         // N x Hn
    const float* b2;
    float* S;                // scratch B x N (spins)
    float* A;                // B x Hn (activations)
    float* dZ;               // B x N
    float* dU;               // B x Hn
    float* loss;             // B
    float* gW1;              // Hn x N
    float* gb1;
    float* gW2;              // N x Hn
    float* gb2;
    float* part;             // ks x G partial gradients (G = gW1 .. gb2, contiguous from gW1)
    int ks;                  // batch slices (train_slices)
    int64_t kc;              // records per slice
};
// batch slices of the gradient reduction: one per 1024 records, at most 256
__host__ __device__ inline int train_slices(int64_t B) {
    const int64_t k = (B + 1023) / 1024;
    return (int)(k < 1 ? 1 : k > 256 ? 256 : k);
}
hipError_t launch_train_grad(const TrainArgs& a, hipStream_t st);

// structured (basic RBM) S.x / moments, bf16 MFMA at FP32 accuracy; see kern_struct.hpp
struct StructArgs {
    int64_t S;
    int N, M, P, ld, nw, n_groups, mode;  // mode 0: S.x, 1: y = E_loc, 2: y = 1
    const uint64_t* states;
    const cplx<float>* T;
    const double* obs;
    const double* x;
    cplx<float>* part;  // nblocks x ld: [z_v | z_h | Z | sum |t|^2 (mode 2)]
    const CgStatus* status;
};
// LDS bytes of k_struct for padded sizes, 0 if no instance covers N, M
size_t struct_lds_bytes(int N, int M);
// resident blocks of k_struct per CU (0 if no instance covers N, M) and the
// waves of one block
int struct_blocks_per_cu(int N, int M, int* waves_per_block);
// e0 / e1 (optional): events recorded at the start and end of the dispatch
hipError_t launch_struct(const StructArgs& a, int blocks, hipStream_t st, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr);
template <typename R>
hipError_t launch_sx_reduce(const SxArgs<R>& a, hipStream_t st);
// column sums of the block partials and the finished product y = S x in one
// pass (k_sx_colsum), ceil(P / kColBlock) blocks (pdot entries)
template <typename R>
hipError_t launch_sx_colsum(const SxArgs<R>& a, hipStream_t st);
constexpr int kColBlock = 64;  // columns per k_sx_colsum block

// explicit S (kern_dense.hip): G = X^T X of the O matrix viewed as S x ncol
// reals (ncol = 2 ld), 64 x 64 upper-triangle tiles, ksplit sample slices
// of kper samples (a multiple of 32) each writing its own partial tiles
template <typename R>
struct SyrkArgs {
    const R* X;
    int64_t S, kper;
    int ncol, nvalid, ntile, npairs, ksplit;  // nvalid = 2 P: columns past it read as 0
    R* part;                                  // ksplit x npairs x 64 x 64
    double* G;                                // npairs x 64 x 64: the slices summed
    const double* wsx;                        // non-null: rows scaled by sqrt(S w_k / sum w)
    const double* mean;                       // non-null: columns centred by <O> (2 P doubles) before the
                                              // products, so S needs no <O><O>^H subtraction
};
// the partial tiles, then G (ncol a multiple of 4; a ksplit that is a
// multiple of 8 keeps each K slice on one XCD)
template <typename R>
hipError_t launch_syrk(const SyrkArgs<R>& a, hipStream_t st);
// S (P x P complex, column-major) from G: mode 0 the CG
// regularisation, mode 1 the direct solver's; reg_dev non-null: r1, r2, rd
// read from device memory instead
template <typename R>
hipError_t launch_s_from_g(const SyrkArgs<R>& a, int P, int n_neural, int64_t S, const double* mean,
                           const double* var, const double* maxvar, double r1, double r2, double rd,
                           const double* reg_dev, int mode, double* Sm, hipStream_t st);
// y = S x for the Hermitian S (column-major): y_i = sum_j conj(S_ji) x_j, so
// wave i reads column i contiguously (the loads of all its column chunks
// issued together); block b (256 threads) takes rows 4 b .. 4 b + 3 and
// writes the partial x^H y of its rows to pdot[2 b] (pdot non-null)
__device__ inline void dense_mv_rows(const double* Sm, const double* x, double* y, int P, double* pdot) {
    __shared__ double red[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = blockIdx.x * 4 + w;
    double yr = 0, yi = 0;
    if (i < P) {
        const double* col = Sm + 2 * (int64_t)i * P;
        for (int j = lane; j < P; j += 64) {
            const double sr = col[2 * j], si = -col[2 * j + 1], xr = x[2 * j], xi = x[2 * j + 1];
            yr += sr * xr - si * xi;
            yi += sr * xi + si * xr;
        }
    }
    yr = group_sum<64>(yr);
    yi = group_sum<64>(yi);
    double dr = 0, di = 0;
    if (i < P) {
        if (lane == 0) {
            y[2 * i] = yr;
            y[2 * i + 1] = yi;
        }
        const double xr = x[2 * i], xi = x[2 * i + 1];
        dr = xr * yr + xi * yi;  // conj(x_i) y_i
        di = xr * yi - xi * yr;
    }
    if (lane == 0) {
        red[0][w] = dr;
        red[1][w] = di;
    }
    __syncthreads();
    if (threadIdx.x == 0 && pdot) {
        pdot[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        pdot[2 * blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}
// y = S x of the Hermitian S, per-block partials of x^H y in pdot
// (ceil(P / 4) blocks); returns at once once st->done
hipError_t launch_dense_mv(const double* Sm, const double* x, double* y, int P, double* pdot, const CgStatus* st,
                           hipStream_t s);

// supported (lanes per chain, units per lane) shapes; wave = 1: the
// one-chain-per-wavefront kernel k_sample_wave (L = 64)
struct Shape {
    int L, K;
    int wave = 0;
};
// (32, 9) and (64, 5) fit the 4 x 72 = 288 units of the 6x6 symmetric
// RBM with alpha = 4 without padding to 512
constexpr Shape kShapes[] = {{8, 1},  {8, 2},  {8, 4},  {8, 8},  {8, 16}, {16, 2},
                             {16, 4}, {16, 8}, {16, 16}, {32, 2}, {32, 4}, {32, 8},
                             {32, 9}, {32, 16}, {64, 1}, {64, 2}, {64, 4}, {64, 5}, {64, 8}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <typename R>
__host__ __device__ inline size_t lds_bytes(const Dims& d) {
    // [C | S | W? | rb | fs? | Pfaffian scratch?] complex R, [val?] complex
    // double, [rowptr? k? group? rowbase? | pb] int32, [bs?] int16,
    // [sites? | mask? | perm] bytes
    size_t b = ((d.w_lds ? 3 : 2) * (size_t)d.N * d.n_alpha + d.N) * sizeof(cplx<R>);
    if (d.pf_rows)
        b += (4 * (size_t)d.pf_rows + (size_t)d.waves * pf_scratch<R>(d.N, d.pf_rows)) * sizeof(cplx<R>) +
             2 * (size_t)d.N * d.N;
    if (d.t_lds) b += (size_t)d.n_ent * 16 + 4 * ((size_t)d.n_rowptr + 3 * d.n_terms) + 8 * (size_t)d.n_terms +
                      d.n_ent;
    b += 8 * (size_t)d.npb;
    b += d.n_symm > 1 ? (size_t)d.n_symm * d.N : 0;
    return (b + 15) & ~(size_t)15;
}

template <typename R>
hipError_t launch_sample(const SampArgs<R>& a, Shape s, int block, hipStream_t st);

// ---- k_sample_wave (FP32, one chain per wavefront, kern_wave.hpp)
// symmetry rows the lane mapping may touch: K (64 / alpha_n) (the padding
// units read copies of row 0 and are masked)
__host__ __device__ inline int wave_ns_pad(const Dims& d, int K) {
    if (d.n_symm == 1) return 1;
    const int r = K * (64 / d.n_alpha);
    return r > d.n_symm ? r : d.n_symm;
}
// entries of the single-site tables (basic RBM: 64 padding entries after
// the last site row for the padding units)
__host__ __device__ inline int wave_n_entries(const Dims& d) { return d.N * d.n_alpha + (d.n_symm == 1 ? 64 : 0); }
// [CS | pair CS] float4, [W | pair W | rb | entry factors] float2, [val]
// complex double, [rowptr k group rowbase pb] int32, [tbond] int16,
// [sites mask perm pair map] bytes (the local-operator tables always in LDS)
__host__ __device__ inline size_t wave_lds_bytes(const Dims& d, int K, int npair) {
    const size_t ne = wave_n_entries(d), np = (size_t)npair * d.n_alpha * 2;
    const int nsp = wave_ns_pad(d, K);
    size_t b = (ne + np) * 16 + (ne + np + d.N + d.n_ent) * 8;
    b += (size_t)d.n_ent * 16 + 4 * ((size_t)d.n_rowptr + 3 * d.n_terms);
    b += 8 * (size_t)d.npb + 2 * (size_t)d.n_terms;
    b += 8 * (size_t)d.n_terms + d.n_ent;
    if (d.n_symm > 1) b += (size_t)nsp * d.N + (npair ? (size_t)nsp * d.npb : 0);
    return (b + 15) & ~(size_t)15;
}
hipError_t launch_sample_wave(const SampArgs<float>& a, int K, int block, hipStream_t st);
// the Pfaffian factor runs one chain per wavefront
inline bool shape_ok_pf(Shape s, int N) { return s.L == 64 && N <= kPfMaxN; }
template <typename R>
hipError_t launch_eval(const EvalArgs<R>& a, Shape s, int block, hipStream_t st);
template <typename R>
hipError_t launch_build_O(const OArgs<R>& a, hipStream_t st);
template <typename R>
hipError_t launch_o_partial(const MomArgs<R>& a, int blocks, hipStream_t st);
// block partials of S.x over the materialised O (k_sx_partial); the caller
// finishes the product with launch_sx_colsum
template <typename R>
hipError_t launch_sx(const SxArgs<R>& a, int blocks, hipStream_t st, hipEvent_t e0 = nullptr,
                     hipEvent_t e1 = nullptr);
// ZM (P <= 64*ZM) supported by launch_sx
int sx_max_params();
// waves per k_sx_partial block for an O row length ld (short rows: one
// block of 16 waves per CU)
inline int sx_waves(int64_t ld) { return ld <= 128 * 4 ? 16 : 4; }

}  // namespace rbmgpu
