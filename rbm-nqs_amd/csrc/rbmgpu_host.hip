// rbmgpu_host.hip -- context, device buffers and the C ABI of include/rbmgpu.h.
//
// One context holds the model tables, the weight tables of the current
// parameters, the sample buffers of one epoch and the CG workspace, all in
// HBM on the current device, and one HIP stream (default priority).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/rbmgpu.h"
#include "kern_api.hpp"

using namespace rbmgpu;
using zd = std::complex<double>;

int rbmgpu::sx_max_params() { return 128 * 24; }

struct rbmgpu_ctx {
    // ---- model
    int N = 0, nvb = 0, na = 0, nsymm = 1, H = 0, nw = 1, pf_rows = 0, pf_refresh = 0;
    int64_t P = 0, P_rbm = 0, ld = 0;  // ld: O row length (P rounded up to even)
    std::vector<int32_t> perm, basis, pb;
    int n_terms = 0, n_groups = 1, n_rowptr = 0, n_ent = 0, tk_max = 0, n_ratio = 0;
    int eloc_bpc = 0;  // resident k_eloc blocks per CU (0: not yet queried)
    // ---- sampler
    int n_chains = 1, warmup = 0, step = 1, prec = 64, lanes_req = 0, keep_o = 1;
    bool auto_chains = false;  // the autotuner picks n_chains (config n_chains = 0)
    int train_dense = -1;      // device SR epoch: 1 CG on the explicit S, 0 matrix-free, -1 undecided
    int chains_cap = 1;        // per-chain buffers hold this many chains
    int64_t n_samples = 0;
    double bond_flips = 0.5;
    uint64_t seed = 0;
    std::vector<zd> params;
    // ---- device
    hipStream_t stream = nullptr;
    int n_cu = 256;
    size_t lds_limit = 65536;
    void *d_W = nullptr, *d_C = nullptr, *d_S = nullptr, *d_hb = nullptr, *d_rb = nullptr,
         *d_lb = nullptr;
    uint8_t* d_perm = nullptr;
    int16_t* d_bs = nullptr;  // Pfaffian index map, bs[i*N + j]
    void* d_fs = nullptr;     // Pfaffian weights in the weight block
    int32_t *d_invperm = nullptr, *d_vbsite = nullptr, *d_pb = nullptr;
    int32_t *d_tk = nullptr, *d_tgroup = nullptr, *d_rowbase = nullptr, *d_rowptr = nullptr;
    int32_t *d_ratio_term = nullptr, *d_term_slot = nullptr;
    uint8_t *d_tsites = nullptr, *d_mask = nullptr;
    double* d_val = nullptr;
    // k_sample_wave (symmetric RBM): ordered image pairs of the proposal bonds
    int npair = 0;
    uint8_t* d_pmap = nullptr;   // nsymm x npb
    int32_t* d_pairs = nullptr;  // npair x 2
    int16_t* d_tbond = nullptr;  // per term: proposal bond of its two sites, -1
    // samples
    int64_t cap = 0, S = 0;
    uint64_t* d_states = nullptr;
    double* d_obs = nullptr;
    void* d_T = nullptr;
    void* d_O = nullptr;
    void* d_Opf = nullptr;  // S x 4 pf_rows Pfaffian derivatives
    uint32_t *d_acc = nullptr, *d_prop = nullptr;
    bool have_chains = false;
    // structured: basic RBM in FP32 without an O matrix (kern_struct.hpp)
    bool structured = false;
    int64_t ld_s = 0;  // structured partial row: [P | sum |t|^2 (M)], even
    // moments
    double *d_part = nullptr, *d_omean = nullptr, *d_ovar = nullptr, *d_eo = nullptr,
           *d_maxvar = nullptr, *d_emom = nullptr, *d_epart = nullptr;
    int64_t epart_cap = 0;  // bytes of d_epart
    double* d_varpart = nullptr;  // chunk partials of the structured Var(O)
    int64_t varpart_cap = 0;
    int64_t part_cap = 0;
    bool moments_valid = false;
    // CG / S.x
    double *d_x = nullptr, *d_b = nullptr, *d_r = nullptr, *d_p = nullptr, *d_Ap = nullptr, *d_pdot = nullptr;
    unsigned* d_cgcnt = nullptr;  // blocks done in a fused dense mat-vec + CG step (0 between launches)
    CgStatus* d_st = nullptr;
    void* d_sxpart = nullptr;
    double* d_sxpart2 = nullptr;
    int64_t sx_blocks = 0, sxpart_cap = 0;
    static constexpr int kSxChunks = 32;  // chunk partials of the moment sums
    // eval scratch
    uint64_t* d_evst = nullptr;
    double *d_evout = nullptr, *d_evobs = nullptr;
    int64_t ev_cap = 0;
    // launch shape
    Shape shape{0, 0};
    int block = 256;
    // timing
    hipEvent_t ev[14] = {};
    double times[6] = {0, 0, 0, 0, 0, 0};
    int64_t cg_last_itn = 0;  // iterations of the last CG solve
    // ---- device-resident SR epoch (rbmgpu_train_epoch)
    double* d_params = nullptr;        // FP64 parameters on the device
    bool params_dev = false;           // d_params is newer than params
    void* d_par = nullptr;             // DevPar
    void* h_par = nullptr;             // pinned host mirror
    void* d_stats = nullptr;           // DevStats
    void* h_stats = nullptr;           // pinned host mirror
    const double* reg_dev = nullptr;   // set while an epoch body is enqueued
    const uint32_t* epoch_dev = nullptr;
    bool capturing = false;
    uint64_t gen = 0;                  // bumped by every device (re)allocation
    uint64_t graph_gen = 0;
    std::vector<std::pair<int64_t, hipGraphExec_t>> graphs;  // CG budget -> instantiated epoch
    int64_t train_itn = 0;             // CG iterations of the last device epoch
    // ---- exact enumeration (rbmgpu_full_epoch): a weighted sample set
    bool weighted = false;
    double *d_w = nullptr, *d_wsx = nullptr;  // w_k = |psi_k|^2 / max, S w_k / sum w
    double* d_lp = nullptr;                    // log psi of the enumerated states
    double* d_wred = nullptr;                  // block partials of the weight reductions
    int64_t w_cap = 0;
    // ---- explicit S (cg-single / direct)
    double* d_Smat = nullptr;  // P x P complex, column-major
    void* d_gpart = nullptr;   // SYRK partial tiles
    int64_t gpart_cap = 0;
    std::vector<char> staging;
    // ---- learned proposal
    int Hn = 0, Np = 0, kmax = 0;
    double lam = 0;
    int64_t learn_t = 0;  // Adam steps
    std::vector<uint8_t> cyc, seedc, seedl;
    uint8_t *d_cyc = nullptr, *d_seedc = nullptr, *d_seedl = nullptr;
    uint16_t *d_W1b = nullptr, *d_W2b = nullptr;  // bf16 copies for the sampler
    float *d_b1p = nullptr, *d_b2p = nullptr;      // biases (b2 padded to Np)
    float* d_mlp = nullptr;    // masters [W1 | b1 | W2 | b2]
    float* d_adam = nullptr;   // Adam moments m, v
    float* d_grad = nullptr;   // gradient in the master layout
    uint64_t* d_rec_st = nullptr;
    int16_t* d_rec_seed = nullptr;
    uint8_t* d_rec_k = nullptr;
    float* d_rec_a = nullptr;
    uint32_t *d_lacc = nullptr, *d_lprop = nullptr;
    float* d_tr = nullptr;     // training scratch
    int64_t tr_cap = 0;
    uint64_t* d_tst = nullptr;
    int16_t* d_tseed = nullptr;
    float* d_tw = nullptr;
    float* d_tpart = nullptr;  // per batch slice partial gradients
    bool learned_epoch = false;  // the last epoch ran the learned sampler
    int64_t mlp_n() const { return (int64_t)Hn * N + Hn + (int64_t)N * Hn + N; }
    // device address of the static scratch array on this context's device
    void* scratch_dev = nullptr;
    std::string err;
};

// Static device scratch of the allocation-free S.x (rbmgpu_apply_S_noalloc):
// the S.x block partials.  Each GPU the library runs on holds one copy
// until the process ends; every call on that GPU shares it.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// device (g_scratch), never through the array's name.
constexpr int64_t kScratchBytes = (int64_t)8 << 20;  // 8 MiB
static __device__ __align__(256) char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch[64] = {};

namespace {

#define HIPCK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);             \
            return -2;                                                             \
        }                                                                          \
    } while (0)

int fail(rbmgpu_ctx* ctx, const std::string& m) {
    ctx->err = m;
    return -1;
}

template <typename T>
int dalloc(rbmgpu_ctx* ctx, T** p, size_t bytes) {
    ctx->gen++;  // captured graphs hold the old addresses
    if (ctx->capturing) return fail(ctx, "device allocation during graph capture");
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (bytes == 0) bytes = 16;
    HIPCK(hipMalloc((void**)p, bytes));
    return 0;
}
template <typename T>
int upload(rbmgpu_ctx* ctx, T** p, const void* src, size_t bytes) {
    if (int rc = dalloc(ctx, p, bytes)) return rc;
    if (bytes) HIPCK(hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

size_t rsize(const rbmgpu_ctx* c) { return c->prec == RBMGPU_FP32 ? 4 : 8; }

// LDS staging: everything if it fits, else drop the local-operator tables,
// then W (both are then read from memory)
Dims dims(const rbmgpu_ctx* c, int block = 256) {
    Dims d{c->N, c->H, c->na, c->nsymm, c->nvb, c->nw, 1, 1, (int)c->pb.size() / 2, c->n_terms, c->n_rowptr,
           c->n_ent, c->pf_rows, block / 64};
    auto fits = [&] {
        return (c->prec == RBMGPU_FP32 ? lds_bytes<float>(d) : lds_bytes<double>(d)) <= c->lds_limit;
    };
    const int order[4][2] = {{1, 1}, {1, 0}, {0, 1}, {0, 0}};
    for (auto& o : order) {
        d.w_lds = o[0];
        d.t_lds = o[1];
        if (fits()) break;
    }
    return d;
}

// weight tables from the FP64 parameter vector (rbm_base.cpp:117-131 layout)
int build_weight_tables(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    int N = c->N, na = c->na, nvb = c->nvb;
    std::vector<zd> W(N * na), C(N * na), S(N * na), hb(na), rb(N), lb(N, 0.);
    for (int k = 0; k < N; k++)
        for (int a = 0; a < na; a++) {
            zd w = c->params[nvb + na + k + (int64_t)a * N];
            W[k * na + a] = w;
            C[k * na + a] = std::cosh(2. * w);
            S[k * na + a] = std::sinh(2. * w);
        }
    for (int a = 0; a < na; a++) hb[a] = c->params[nvb + a];
    for (int f = 0; f < N; f++) rb[f] = c->params[f % nvb];
    for (int i = 0; i < nvb; i++)
        for (int s = 0; s < c->nsymm; s++) lb[c->perm[s * N + c->basis[i]]] += c->params[i];
    // one staging block [W | C | S | hb | rb | lb] copied asynchronously on
    // the context stream; the staging buffer lives in the context so the
    // copy may complete after this call returns
    const size_t nwt = (size_t)N * na, nfs = 4 * (size_t)c->pf_rows,
                 total = 3 * nwt + na + 2 * (size_t)N + nfs;
    const size_t rs = rsize(c) * 2;
    if (!c->d_W) {
        char* base = nullptr;
        HIPCK(hipMalloc((void**)&base, total * rs));
        c->d_W = base;
        c->d_C = base + nwt * rs;
        c->d_S = base + 2 * nwt * rs;
        c->d_hb = base + 3 * nwt * rs;
        c->d_rb = base + (3 * nwt + na) * rs;
        c->d_lb = base + (3 * nwt + na + N) * rs;
        c->d_fs = base + (3 * nwt + na + 2 * N) * rs;
    }
    HIPCK(hipStreamSynchronize(c->stream));  // previous copy out of the staging block
    c->staging.resize(total * rs);
    auto pack = [&](size_t off, const std::vector<zd>& v) {
        for (size_t i = 0; i < v.size(); i++) {
            if (c->prec == RBMGPU_FP64) {
                reinterpret_cast<zd*>(c->staging.data())[off + i] = v[i];
            } else {
                reinterpret_cast<std::complex<float>*>(c->staging.data())[off + i] = std::complex<float>(v[i]);
            }
        }
    };
    pack(0, W);
    pack(nwt, C);
    pack(2 * nwt, S);
    pack(3 * nwt, hb);
    pack(3 * nwt + na, rb);
    pack(3 * nwt + na + N, lb);
    if (nfs) {  // fs_(r, c) at parameter P_rbm + r + c * rows -> fs[r*4 + c]
        std::vector<zd> fs(nfs);
        for (int r = 0; r < c->pf_rows; r++)
            for (int q = 0; q < 4; q++) fs[r * 4 + q] = c->params[c->P_rbm + r + (int64_t)q * c->pf_rows];
        pack(3 * nwt + na + 2 * N, fs);
    }
    HIPCK(hipMemcpyAsync(c->d_W, c->staging.data(), total * rs, hipMemcpyHostToDevice, c->stream));
    c->moments_valid = false;
    return 0;
}

template <typename R>
Tabs<R> tabs(const rbmgpu_ctx* c) {
    return {(const cplx<R>*)c->d_W,  (const cplx<R>*)c->d_C,  (const cplx<R>*)c->d_S,
            (const cplx<R>*)c->d_hb, (const cplx<R>*)c->d_rb, (const cplx<R>*)c->d_lb,
            c->d_perm,               (const cplx<R>*)c->d_fs, c->d_bs};
}
Terms terms(const rbmgpu_ctx* c) {
    return {c->n_terms, c->n_groups, c->d_tk,     c->d_tsites, c->d_tgroup,
            c->d_rowbase, c->d_rowptr, c->d_mask, c->d_val};
}

int ensure_samples(rbmgpu_ctx* c, int64_t S) {
    rbmgpu_ctx* ctx = c;
    if (S <= c->cap) return 0;
    if (int rc = dalloc(ctx, &c->d_states, S * c->nw * sizeof(uint64_t))) return rc;
    if (int rc = dalloc(ctx, &c->d_obs, S * 2 * c->n_groups * sizeof(double))) return rc;
    if (int rc = dalloc(ctx, &c->d_T, S * c->H * 2 * rsize(c))) return rc;
    if (!c->structured)
        if (int rc = dalloc(ctx, &c->d_O, S * c->ld * 2 * rsize(c))) return rc;
    if (c->pf_rows)
        if (int rc = dalloc(ctx, &c->d_Opf, S * 4 * c->pf_rows * 2 * rsize(c))) return rc;
    if (c->Hn) {
        if (int rc = dalloc(ctx, &c->d_rec_st, S * c->nw * sizeof(uint64_t))) return rc;
        if (int rc = dalloc(ctx, &c->d_rec_seed, S * sizeof(int16_t))) return rc;
        if (int rc = dalloc(ctx, &c->d_rec_k, S)) return rc;
        if (int rc = dalloc(ctx, &c->d_rec_a, S * sizeof(float))) return rc;
    }
    c->cap = S;
    return 0;
}

bool shape_fits(const rbmgpu_ctx* c, Shape s) {
    return s.L * s.K >= c->H && (c->pf_rows == 0 || shape_ok_pf(s, c->N));
}

// block size: 256 threads unless that leaves fewer than two blocks per CU;
// with the Pfaffian also small enough for its per-wave LDS scratch
int pick_block(const rbmgpu_ctx* c, int L, int64_t groups) {
    int b = 256;
    while (b > 64 && groups * L / b < 2 * c->n_cu) b /= 2;
    b = std::max(b, L);
    if (c->pf_rows)
        while (b > 64 && (c->prec == RBMGPU_FP32 ? lds_bytes<float>(dims(c, b)) : lds_bytes<double>(dims(c, b))) >
                             c->lds_limit)
            b /= 2;
    return b;
}

template <typename R>
SampArgs<R> samp_args(rbmgpu_ctx* c, uint32_t epoch, int block, int64_t n_samples, int warmup) {
    SampArgs<R> a{};
    a.d = dims(c, block);
    a.t = tabs<R>(c);
    a.T = terms(c);
    a.n_chains = c->n_chains;
    a.n_samples = n_samples;
    a.warmup = warmup;
    a.step = c->step;
    a.npb = (int)c->pb.size() / 2;
    double thr = std::floor(c->bond_flips * 4294967296.0);
    a.thr = (uint64_t)std::min(std::max(thr, 0.0), 4294967296.0);
    a.epoch = epoch;
    a.k0 = (uint32_t)c->seed;
    a.k1 = (uint32_t)(c->seed >> 32);
    a.pb = c->d_pb;
    a.states = c->d_states;
    a.obs = c->d_obs;
    a.Tout = (cplx<R>*)c->d_T;
    a.Opf = (cplx<R>*)c->d_Opf;
    a.pf_refresh = c->pf_refresh;
    a.acc = c->d_acc;
    a.prop = c->d_prop;
    a.pmap = c->d_pmap;
    a.pairs = c->d_pairs;
    a.npair = c->npair;
    a.tbond = c->d_tbond;
    return a;
}

// kept samples' local observables by k_eloc after the sampler (FP32 basic
// RBM, terms of at most 4 sites, tables in LDS) instead of inside it
bool eloc_after(const rbmgpu_ctx* c) {
    return c->prec == RBMGPU_FP32 && c->pf_rows == 0 && c->nsymm == 1 && c->n_ratio >= 1 &&
           c->n_ratio <= kElocThreads && c->tk_max <= 4 && c->nw <= 2 && eloc_lds_bytes(c->N, c->H) <= 65536;
}

// the one-chain-per-wavefront sampler (FP32, no Pfaffian factor, at most
// 128 sites and 512 hidden units; symmetric RBM: alpha divides 64)
bool wave_ok(const rbmgpu_ctx* c) {
    return c->prec == RBMGPU_FP32 && c->pf_rows == 0 && c->nw <= 2 && c->H <= 512 &&
           (c->nsymm == 1 || 64 % c->na == 0) && (int)c->pb.size() / 2 < 32768;
}
// its dims (all tables in LDS)
Dims wave_dims(const rbmgpu_ctx* c, int block, int K, int npair) {
    (void)K;
    (void)npair;
    Dims d = dims(c, block);
    d.w_lds = 1;
    d.t_lds = 1;
    return d;
}

template <typename R>
int run_sample(rbmgpu_ctx* c, uint32_t epoch, Shape s, int block, int64_t n_samples, int warmup) {
    rbmgpu_ctx* ctx = c;
    SampArgs<R> a = samp_args<R>(c, epoch, block, n_samples, warmup);
    a.epoch_dev = c->epoch_dev;
    a.obs_later = eloc_after(c);
    if constexpr (sizeof(R) == 4) {
        if (s.wave) {
            a.npair = c->nsymm > 1 ? c->npair : 0;
            a.d = wave_dims(c, block, s.K, a.npair);
            if (wave_lds_bytes(a.d, s.K, a.npair) > c->lds_limit)
                return fail(c, "the one-chain-per-wavefront sampler's tables do not fit LDS");
            HIPCK(launch_sample_wave(a, s.K, block, c->stream));
            return 0;
        }
    }
    HIPCK(launch_sample<R>(a, s, block, c->stream));
    return 0;
}

int run_eloc(rbmgpu_ctx* c, int64_t n) {
    rbmgpu_ctx* ctx = c;
    ElocArgs a{};
    a.n = n;
    a.N = c->N;
    a.H = c->H;
    a.nw = c->nw;
    a.T = terms(c);
    Tabs<float> t = tabs<float>(c);
    a.C = t.C;
    a.S = t.S;
    a.rb = t.rb;
    a.states = c->d_states;
    a.tanh = (const cplx<float>*)c->d_T;
    a.obs = c->d_obs;
    a.n_ratio = c->n_ratio;
    a.ratio_term = c->d_ratio_term;
    a.term_slot = c->d_term_slot;
    const int64_t per = kElocThreads / c->n_ratio;
    if (!c->eloc_bpc) c->eloc_bpc = eloc_blocks_per_cu(c->N, c->H);
    const int64_t blocks = std::min<int64_t>((n + per - 1) / per, (int64_t)c->eloc_bpc * c->n_cu);
    HIPCK(launch_eloc(a, (int)blocks, c->stream));
    return 0;
}

// ---- learned proposal
LearnArgs learn_args(rbmgpu_ctx* c, uint32_t epoch) {
    LearnArgs a{};
    a.s = samp_args<float>(c, epoch, 64 * kMlpChains, c->n_samples, c->warmup);
    a.Hn = c->Hn;
    a.Np = c->Np;
    a.kmax = c->kmax;
    double thr = std::floor(c->lam * 4294967296.0);
    a.lam_thr = (uint32_t)std::min(std::max(thr, 0.0), 4294967295.0);
    a.W1 = c->d_W1b;
    a.W2 = c->d_W2b;
    a.b1 = c->d_b1p;
    a.b2 = c->d_b2p;
    a.cyc = c->d_cyc;
    a.seedc = c->d_seedc;
    a.seedl = c->d_seedl;
    a.rec_st = c->d_rec_st;
    a.rec_seed = c->d_rec_seed;
    a.rec_k = c->d_rec_k;
    a.rec_a = c->d_rec_a;
    a.lacc = c->d_lacc;
    a.lprop = c->d_lprop;
    return a;
}

// smallest units-per-lane K of the learned kernel with 64 K >= H
int learn_k(const rbmgpu_ctx* c) {
    for (int k : {1, 2, 4, 5, 8})
        if (64 * k >= c->H) return k;
    return 0;
}

// bf16 copies and padded biases from the FP32 masters
__global__ void k_mlp_pack(const float* mlp, int N, int Np, int Hn, uint16_t* W1b, uint16_t* W2b, float* b1p,
                           float* b2p) {
    const float* W1 = mlp;
    const float* b1 = W1 + (int64_t)Hn * N;
    const float* W2 = b1 + Hn;
    const float* b2 = W2 + (int64_t)N * Hn;
    auto bf = [](float f) {
        uint32_t b = __float_as_uint(f);
        b += 0x7FFFu + ((b >> 16) & 1u);
        return (uint16_t)(b >> 16);
    };
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)Hn * Np) {
        int h = (int)(i / Np), j = (int)(i % Np);
        W1b[i] = j < N ? bf(W1[(int64_t)h * N + j]) : 0;
        int o = (int)(i / Hn), hh = (int)(i % Hn);
        W2b[i] = o < N ? bf(W2[(int64_t)o * Hn + hh]) : 0;
    }
    if (i < Hn) b1p[i] = b1[i];
    if (i < Np) b2p[i] = i < N ? b2[i] : 0.f;
}

// Adam (Kingma & Ba) on the FP32 masters
__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float c1, float c2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[i];
    float mi = b1 * m[i] + (1 - b1) * gi, vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int learn_pack(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    int64_t n = std::max<int64_t>((int64_t)c->Hn * c->Np, std::max(c->Hn, c->Np));
    hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_mlp, c->N,
                       c->Np, c->Hn, c->d_W1b, c->d_W2b, c->d_b1p, c->d_b2p);
    HIPCK(hipGetLastError());
    return 0;
}

int ensure_train(rbmgpu_ctx* c, int64_t B) {
    rbmgpu_ctx* ctx = c;
    if (B <= c->tr_cap) return 0;
    // S (B x N), A (B x Hn), dZ (B x N), dU (B x Hn), loss (B), w (B)
    if (int rc = dalloc(ctx, &c->d_tr, (size_t)B * (2 * c->N + 2 * c->Hn + 1) * sizeof(float))) return rc;
    if (int rc = dalloc(ctx, &c->d_tst, (size_t)B * c->nw * sizeof(uint64_t))) return rc;
    if (int rc = dalloc(ctx, &c->d_tseed, (size_t)B * sizeof(int16_t))) return rc;
    if (int rc = dalloc(ctx, &c->d_tw, (size_t)B * sizeof(float))) return rc;
    if (int rc = dalloc(ctx, &c->d_tpart, (size_t)train_slices(B) * c->mlp_n() * sizeof(float))) return rc;
    c->tr_cap = B;
    return 0;
}

// gradient of the batch objective into d_grad; loss to *loss
int learn_grad(rbmgpu_ctx* c, int64_t B, double beta, double* loss) {
    rbmgpu_ctx* ctx = c;
    TrainArgs t{};
    t.N = c->N;
    t.Np = c->Np;
    t.Hn = c->Hn;
    t.nw = c->nw;
    t.B = B;
    t.st = c->d_tst;
    t.seed = c->d_tseed;
    t.w = c->d_tw;
    t.beta = (float)beta;
    t.inv_b = 1.f / (float)B;
    t.W1 = c->d_mlp;
    t.b1 = t.W1 + (int64_t)c->Hn * c->N;
    t.W2 = t.b1 + c->Hn;
    t.b2 = t.W2 + (int64_t)c->N * c->Hn;
    t.S = c->d_tr;
    t.A = t.S + B * c->N;
    t.dZ = t.A + B * c->Hn;
    t.dU = t.dZ + B * c->N;
    t.loss = t.dU + B * c->Hn;
    t.gW1 = c->d_grad;
    t.gb1 = t.gW1 + (int64_t)c->Hn * c->N;
    t.gW2 = t.gb1 + c->Hn;
    t.gb2 = t.gW2 + (int64_t)c->N * c->Hn;
    t.part = c->d_tpart;
    t.ks = train_slices(B);
    t.kc = (B + t.ks - 1) / t.ks;
    HIPCK(launch_train_grad(t, c->stream));
    std::vector<float> lv(B);
    HIPCK(hipMemcpyAsync(lv.data(), t.loss, B * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    double L = 0;
    for (float x : lv) L += x;
    if (loss) *loss = L / (double)B;
    return 0;
}

// cluster catalogue tables: cyc (N x kmax) and, per (i, k), the seeds j
// whose first-k clusters are the same site set
int learn_tables(rbmgpu_ctx* c, const int32_t* clusters) {
    const int N = c->N, K = c->kmax;
    c->cyc.assign((size_t)N * K, 0);
    c->seedc.assign((size_t)N * K, 0);
    c->seedl.assign((size_t)N * K * 6, 0);
    for (int i = 0; i < N; i++)
        for (int t = 0; t < K; t++) {
            int v = clusters[i * K + t];
            if (v < 0 || v >= N) return fail(c, "cluster site out of range");
            c->cyc[i * K + t] = (uint8_t)v;
        }
    for (int i = 0; i < N; i++)
        if (c->cyc[i * K] != i) return fail(c, "cluster of site i must start at i");
    for (int i = 0; i < N; i++)
        for (int k = 1; k <= K; k++) {
            std::vector<int> a(c->cyc.begin() + i * K, c->cyc.begin() + i * K + k);
            std::sort(a.begin(), a.end());
            int cnt = 0;
            for (int j = 0; j < N; j++) {
                std::vector<int> b(c->cyc.begin() + j * K, c->cyc.begin() + j * K + k);
                std::sort(b.begin(), b.end());
                if (a == b) {
                    if (cnt == 6) return fail(c, "more than 6 seeds of one cluster");
                    c->seedl[(i * K + k - 1) * 6 + cnt++] = (uint8_t)j;
                }
            }
            c->seedc[i * K + k - 1] = (uint8_t)cnt;
        }
    return 0;
}

template <typename R>
int run_build_O(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (c->structured || c->S == 0) return 0;
    OArgs<R> a{};
    a.d = dims(c);
    a.S = c->S;
    a.P = (int)c->P;
    a.P_rbm = (int)c->P_rbm;
    a.ld = (int)c->ld;
    a.states = c->d_states;
    a.T = (const cplx<R>*)c->d_T;
    a.vb_site = c->d_vbsite;
    a.invperm = c->d_invperm;
    a.Opf = c->pf_rows ? (const cplx<R>*)c->d_Opf : nullptr;
    a.O = (cplx<R>*)c->d_O;
    HIPCK(launch_build_O<R>(a, c->stream));
    return 0;
}

// automatic chain count (config n_chains = 0): for chain counts 2^k from
// 1024 to min(n_samples, 65536), time a short epoch (4 warm-up sweeps + 2
// kept) with the tuned launch shape, and keep the count that minimises the
// predicted epoch, per-sweep time x (warm-up + ceil(n_samples / chains) x
// sweeps per sample)
int tune_chains(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (!c->auto_chains) return 0;
    std::vector<int> cand;
    for (int64_t k = 1024; k <= c->chains_cap; k *= 2) cand.push_back((int)k);
    if (cand.empty() || cand.back() != c->chains_cap) cand.push_back(c->chains_cap);
    if (int rc = ensure_samples(c, 2 * (int64_t)c->chains_cap)) return rc;
    const int keep = c->n_chains;
    double best = 1e300;
    int best_c = keep;
    for (int nc : cand) {
        c->n_chains = nc;
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {
            HIPCK(hipEventRecord(c->ev[10], c->stream));
            const int64_t ns = 2 * (int64_t)nc;
            int rc = c->prec == RBMGPU_FP32 ? run_sample<float>(c, 0xFFFFFFF1u, c->shape, c->block, ns, 4)
                                            : run_sample<double>(c, 0xFFFFFFF1u, c->shape, c->block, ns, 4);
            if (rc) {
                c->n_chains = keep;
                return rc;
            }
            HIPCK(hipEventRecord(c->ev[11], c->stream));
            HIPCK(hipEventSynchronize(c->ev[11]));
            HIPCK(hipEventElapsedTime(&ms, c->ev[10], c->ev[11]));
        }
        const double sweeps = c->warmup + (double)((c->n_samples + nc - 1) / nc) * c->step;
        const double pred = ms / (4.0 + 2.0 * c->step) * sweeps;
        if (pred < best) {
            best = pred;
            best_c = nc;
        }
    }
    c->n_chains = best_c;
    return 0;
}

int autotune(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (c->shape.L) return 0;
    std::vector<Shape> cand;
    for (int L : {8, 16, 32, 64}) {
        Shape best{0, 0};
        for (int i = 0; i < kNumShapes; i++) {
            Shape s = kShapes[i];
            if (s.L == L && shape_fits(c, s) && (best.L == 0 || s.K < best.K)) best = s;
        }
        if (best.L && (c->lanes_req == 0 || c->lanes_req == L)) cand.push_back(best);
    }
    // the one-chain-per-wavefront sampler, at 4 to 16 chains per block
    // (lanes_req -1: only it)
    std::vector<int> cand_block(cand.size(), 0);
    if (wave_ok(c) && (c->lanes_req == 0 || c->lanes_req == -1)) {
        if (c->lanes_req == -1) {
            cand.clear();
            cand_block.clear();
        }
        const int K = (c->H + 63) / 64, np = c->nsymm > 1 ? c->npair : 0;
        for (int b : {256, 384, 512, 640, 1024}) {
            if (wave_lds_bytes(wave_dims(c, b, K, np), K, np) > c->lds_limit) continue;
            Shape w{64, K};
            w.wave = 1;
            cand.push_back(w);
            cand_block.push_back(b);
        }
    }
    if (cand.empty()) return fail(c, "no launch shape for this number of hidden units");
    if (cand.size() == 1) {
        c->shape = cand[0];
        c->block = cand_block[0] ? cand_block[0] : pick_block(c, c->shape.L, c->n_chains);
        return tune_chains(c);
    }
    // time a short epoch per shape: 2 samples per chain after up to 30
    // warm-up sweeps, so the per-block table staging and the kept samples'
    // local observables weigh about as much as in the real epoch
    if (int rc = ensure_samples(c, 2 * (int64_t)c->n_chains)) return rc;
    const int trial_warm = std::max(2, std::min(c->warmup, 30));
    float best_ms = 1e30f;
    for (size_t ci = 0; ci < cand.size(); ci++) {
        const Shape s = cand[ci];
        const int block = cand_block[ci] ? cand_block[ci] : pick_block(c, s.L, c->n_chains);
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {
            HIPCK(hipEventRecord(c->ev[10], c->stream));
            int64_t ns = 2 * (int64_t)c->n_chains;
            int rc = c->prec == RBMGPU_FP32 ? run_sample<float>(c, 0xFFFFFFF0u, s, block, ns, trial_warm)
                                            : run_sample<double>(c, 0xFFFFFFF0u, s, block, ns, trial_warm);
            if (rc) return rc;
            HIPCK(hipEventRecord(c->ev[11], c->stream));
            HIPCK(hipEventSynchronize(c->ev[11]));
            HIPCK(hipEventElapsedTime(&ms, c->ev[10], c->ev[11]));
        }
        if (ms < best_ms) {
            best_ms = ms;
            c->shape = s;
            c->block = block;
        }
    }
    return tune_chains(c);
}

// zero fill by a kernel: unlike hipMemsetAsync it may be captured into a
// HIP graph and replayed
__global__ void __launch_bounds__(256) k_zero(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
int zero_async(rbmgpu_ctx* c, void* p, size_t bytes) {
    rbmgpu_ctx* ctx = c;
    const int64_t n = (int64_t)(bytes / 4);
    if (n == 0) return 0;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4 * (int64_t)c->n_cu);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, c->stream, (uint32_t*)p, n);
    HIPCK(hipGetLastError());
    return 0;
}
// timing events are left out of captured graphs
int record(rbmgpu_ctx* c, int e) {
    rbmgpu_ctx* ctx = c;
    if (!c->capturing) HIPCK(hipEventRecord(c->ev[e], c->stream));
    return 0;
}

float elapsed(rbmgpu_ctx* c, int a, int b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0;
    return ms;
}

// ---- reductions in FP64 (aggregator.cpp:84-239, one rank, unit weights)
// ---- CG on the device, FP64 vectors (cg_solver.cpp:119-189)
// block-wide sums of x and y, returned to every thread (blockDim a multiple
// of 64, at most 1024): wave shuffles, then one LDS pass over the waves
__device__ void block_sum2(double& x, double& y) {
    __shared__ double ra[16], rb[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    x = group_sum<64>(x);
    y = group_sum<64>(y);
    if (lane == 0) {
        ra[w] = x;
        rb[w] = y;
    }
    __syncthreads();
    x = y = 0;
    for (int q = 0; q < nw; q++) {
        x += ra[q];
        y += rb[q];
    }
    __syncthreads();  // ra / rb may be written again by the next call
}

// <O>, Var(O), <E O*> per parameter from k_o_partial's chunks: block p,
// its threads over the chunks, block sums in a fixed order.  Chunk b
// (weight W_b = sum w, n_b samples without weights): mean_b = K_b +
// sum w d / W_b, M2_b = sum w |d|^2 - |sum w d|^2 / W_b; merged as M2 =
// sum_b M2_b + W_b |mean_b - mean|^2 (Chan, Golub and LeVeque; the
// reference accumulates weighted Welford updates, aggregator.cpp:190-231),
// divided by W - sum w^2 / W (aggregator.cpp:141-154)
__global__ void __launch_bounds__(256) k_o_final(const double* part, int64_t nb, int P, double* mean, double* var,
                                                 double* eo, unsigned long long* maxvar) {
    const int p = blockIdx.x;
    double sr = 0, si = 0, er = 0, ei = 0, W = 0, W2 = 0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) {
        const double* q = part + (b * P + p) * kOPartRow;
        sr += q[7] * q[0] + q[2];
        si += q[7] * q[1] + q[3];
        er += q[5];
        ei += q[6];
        W += q[7];
        W2 += q[8];
    }
    block_sum2(sr, si);
    block_sum2(er, ei);
    block_sum2(W, W2);
    const double inv = 1.0 / W;
    const double mr = sr * inv, mi = si * inv;
    double m2 = 0, dummy = 0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) {
        const double* q = part + (b * P + p) * kOPartRow;
        const double n = q[7];
        if (n == 0) continue;
        const double dr = q[2] / n, di = q[3] / n;              // mean_b - K_b
        const double ur = q[0] + dr - mr, ui = q[1] + di - mi;  // mean_b - mean
        m2 += (q[4] - n * (dr * dr + di * di)) + n * (ur * ur + ui * ui);
    }
    block_sum2(m2, dummy);
    if (threadIdx.x == 0) {
        mean[2 * p] = mr;
        mean[2 * p + 1] = mi;
        const double den = W - W2 / W;
        const double v = den > 0 ? fmax(m2 / den, 0.0) : 0.0;
        var[p] = v;
        eo[2 * p] = er * inv;
        eo[2 * p + 1] = ei * inv;
        atomicMax(maxvar, (unsigned long long)__double_as_longlong(v));
    }
}

// Energy moments (aggregator.cpp:84-154 with binning; weights w of an
// enumerated set, 1 otherwise) in two launches.  k_e_part: one block per
// range of `chunk` samples (a range never straddles a bin): the range's
// weighted sum of every group, and for the energy (group 0) the weighted
// sum of squared deviations from the range's own weighted mean.
// part row: [M2, W, (sum w x re, im) x n_groups, sum w^2].
__global__ void __launch_bounds__(256) k_e_part(const double* obs, const double* w, int64_t S, int n_groups,
                                                int64_t chunk, double* part) {
    const int G = 2 * n_groups;
    const int64_t k0 = (int64_t)blockIdx.x * chunk, k1 = k0 + chunk < S ? k0 + chunk : S;
    double* out = part + (int64_t)blockIdx.x * (3 + G);
    double sw = 0, sw2 = 0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const double wk = w ? w[k] : 1.0;
        sw += wk;
        sw2 += wk * wk;
    }
    block_sum2(sw, sw2);
    double mr = 0, mi = 0;
    for (int g = 0; g < n_groups; g++) {
        double sr = 0, si = 0;
        for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
            const double wk = w ? w[k] : 1.0;
            sr += wk * obs[k * G + 2 * g];
            si += wk * obs[k * G + 2 * g + 1];
        }
        block_sum2(sr, si);
        if (g == 0 && sw > 0) {
            mr = sr / sw;
            mi = si / sw;
        }
        if (threadIdx.x == 0) {
            out[2 + 2 * g] = sr;
            out[3 + 2 * g] = si;
        }
    }
    double v = 0, dummy = 0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const double wk = w ? w[k] : 1.0;
        const double dr = obs[k * G] - mr, di = obs[k * G + 1] - mi;
        v += wk * (dr * dr + di * di);
    }
    block_sum2(v, dummy);
    if (threadIdx.x == 0) {
        out[0] = v;
        out[1] = sw;
        out[2 + G] = sw2;
    }
}

// k_e_final, one block: the means from the range sums, the energy's sum of
// squared deviations by the pairwise merge M2 = sum_b (M2_b + W_b |m_b - m|^2)
// (Chan, Golub and LeVeque; the reference merges its ranks' partial results
// the same way, aggregator.cpp:101-116), the unbiased factor W - sum w^2 / W
// (aggregator.cpp:141-154), tau as aggregator.cpp:120-131 (bin_size in
// samples), and the binned variance from the bin sums (per_bin ranges per
// bin).  out: [mean re, mean im, var, var_binned, tau, std, ..., group means
// from 16].
__global__ void __launch_bounds__(1024) k_e_final(const double* part, int64_t nblk, int64_t per_bin, int64_t S,
                                                  int n_groups, int64_t n_bins, double* out) {
    const int PS = 3 + 2 * n_groups;
    double W = 0, W2 = 0;
    for (int64_t b = threadIdx.x; b < nblk; b += blockDim.x) {
        W += part[b * PS + 1];
        W2 += part[b * PS + PS - 1];
    }
    block_sum2(W, W2);
    double mr = 0, mi = 0;
    for (int g = 0; g < n_groups; g++) {
        double sr = 0, si = 0;
        for (int64_t b = threadIdx.x; b < nblk; b += blockDim.x) {
            sr += part[b * PS + 2 + 2 * g];
            si += part[b * PS + 3 + 2 * g];
        }
        block_sum2(sr, si);
        sr /= W;
        si /= W;
        if (g == 0) {
            mr = sr;
            mi = si;
        }
        if (threadIdx.x == 0) {
            out[16 + 2 * g] = sr;
            out[16 + 2 * g + 1] = si;
        }
    }
    double v = 0, dummy = 0;
    for (int64_t b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double n = part[b * PS + 1];
        if (n == 0) continue;
        const double dr = part[b * PS + 2] / n - mr, di = part[b * PS + 3] / n - mi;
        v += part[b * PS] + n * (dr * dr + di * di);
    }
    block_sum2(v, dummy);
    double vb = 0;
    const int64_t bs = n_bins > 1 ? S / n_bins : 0;
    if (n_bins > 1) {
        for (int64_t b = threadIdx.x; b < n_bins; b += blockDim.x) {
            double br = 0, bi = 0, bw = 0;
            for (int64_t j = 0; j < per_bin; j++) {
                br += part[(b * per_bin + j) * PS + 2];
                bi += part[(b * per_bin + j) * PS + 3];
                bw += part[(b * per_bin + j) * PS + 1];
            }
            if (bw > 0) {
                const double dr = br / bw - mr, di = bi / bw - mi;
                vb += bw * (dr * dr + di * di);
            }
        }
        block_sum2(vb, dummy);
    }
    if (threadIdx.x == 0) {
        double denom = W - W2 / W;  // ptotal - wsum2 / ptotal
        double tau = 0;
        if (n_bins > 1 && !(v == 0 && vb == 0)) tau = 0.5 * (fmax(bs * vb / v, 1.0) - 1);
        double var = denom > 0 ? fmax(v / denom, 0.0) : 0.0, varb = denom > 0 ? fmax(vb / denom, 0.0) : 0.0;
        out[0] = mr;
        out[1] = mi;
        out[2] = var;
        out[3] = varb;
        out[4] = tau;
        out[5] = n_bins > 1 ? sqrt(varb / (double)n_bins) : sqrt(var / (double)S);
    }
}

// parameters of a device-resident SR epoch, written by the host before each
// graph replay (abstract_optimizer.cpp:80-178 with the decay of
// abstract_optimizer.cpp:39-48 applied on the host)
struct DevPar {
    uint32_t epoch, pad;
    int64_t max_iter;
    double reg[3];  // r1, r2, r1 delta (read by k_sx_colsum)
    double lr, rtol;
};
// results of a device-resident SR epoch
struct DevStats {
    double e[6];  // k_e_final: mean re/im, var, var binned, tau, std
    double acc;   // mean acceptance over the chains
    double rs, max_dw;
    int64_t itn;
    int32_t ok, pending;
};

__global__ void __launch_bounds__(1024) k_cg_init(const double* b, const double* Ax, double* r, double* p,
                                                  int P, CgStatus* st, int64_t max_iter, double rtol,
                                                  const DevPar* par) {
    if (par) {
        max_iter = par->max_iter;
        rtol = par->rtol;
    }
    double s = 0, dummy = 0;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        double rr = b[2 * i] - Ax[2 * i], ri = b[2 * i + 1] - Ax[2 * i + 1];
        r[2 * i] = rr;
        r[2 * i + 1] = ri;
        p[2 * i] = rr;
        p[2 * i + 1] = ri;
        s += rr * rr + ri * ri;
    }
    block_sum2(s, dummy);
    if (threadIdx.x == 0) {
        st->rsold = s;
        st->rs = s;
        st->it = 0;
        st->itn = max_iter;
        st->max_iter = max_iter;
        st->rtol = rtol;
        st->done = max_iter <= 0;
    }
}

// CG update after y = S p (cg_solver.cpp:160-189), one block: alpha =
// rs / p^H Ap from the column blocks' partials (k_sx_colsum), x += alpha p,
// r -= alpha Ap, |r|^2, the convergence test, beta, p = r + beta p, and the
// status update.  The first kCgReg x blockDim elements stay in registers and
// all their loads are issued before the first reduction, so the step waits
// on global memory once; larger P walks the rest twice.
constexpr int kCgReg = 4;
// one CG step after y = A p (cg_solver.cpp:160-185): alpha from the nb
// partials of p^H Ap, x += alpha p, r -= alpha Ap, convergence on |r|^2,
// p = r + beta p; one block, any multiple of 64 threads
__device__ void cg_update_body(const double* pdot, int nb, double* x, double* r, double* p, const double* Ap, int P,
                               CgStatus* st) {
    const int nt = blockDim.x;
    double pv[kCgReg][2], qv[kCgReg][2], xv[kCgReg][2], rv[kCgReg][2];
#pragma unroll
    for (int k = 0; k < kCgReg; k++) {
        const int i = threadIdx.x + k * nt < P ? threadIdx.x + k * nt : 0;  // P >= 1
        pv[k][0] = p[2 * i];
        pv[k][1] = p[2 * i + 1];
        qv[k][0] = Ap[2 * i];
        qv[k][1] = Ap[2 * i + 1];
        xv[k][0] = x[2 * i];
        xv[k][1] = x[2 * i + 1];
        rv[k][0] = r[2 * i];
        rv[k][1] = r[2 * i + 1];
    }
    const double rsold = st->rsold, rtol = st->rtol;
    double ar = 0, ai = 0;  // p^H Ap
    for (int b = threadIdx.x; b < nb; b += nt) {
        ar += pdot[2 * b];
        ai += pdot[2 * b + 1];
    }
    block_sum2(ar, ai);
    const double den = ar * ar + ai * ai;
    const double alr = rsold * ar / den, ali = -rsold * ai / den;  // rsold / (ar + i ai)
    double sr = 0, dummy = 0;
#pragma unroll
    for (int k = 0; k < kCgReg; k++) {
        const int i = threadIdx.x + k * nt;
        if (i < P) {
            const double pr = pv[k][0], pi = pv[k][1], qr = qv[k][0], qi = qv[k][1];
            x[2 * i] = xv[k][0] + (alr * pr - ali * pi);
            x[2 * i + 1] = xv[k][1] + (alr * pi + ali * pr);
            const double rr = rv[k][0] - (alr * qr - ali * qi), ri = rv[k][1] - (alr * qi + ali * qr);
            rv[k][0] = rr;
            rv[k][1] = ri;
            r[2 * i] = rr;
            r[2 * i + 1] = ri;
            sr += rr * rr + ri * ri;
        }
    }
    for (int i = threadIdx.x + kCgReg * nt; i < P; i += nt) {
        const double pr = p[2 * i], pi = p[2 * i + 1], qr = Ap[2 * i], qi = Ap[2 * i + 1];
        x[2 * i] += alr * pr - ali * pi;
        x[2 * i + 1] += alr * pi + ali * pr;
        const double rr = r[2 * i] - (alr * qr - ali * qi), ri = r[2 * i + 1] - (alr * qi + ali * qr);
        r[2 * i] = rr;
        r[2 * i + 1] = ri;
        sr += rr * rr + ri * ri;
    }
    block_sum2(sr, dummy);
    const bool conv = sr < rtol;
    if (!conv) {
        const double beta = sr / rsold;
#pragma unroll
        for (int k = 0; k < kCgReg; k++) {
            const int i = threadIdx.x + k * nt;
            if (i < P) {
                p[2 * i] = rv[k][0] + beta * pv[k][0];
                p[2 * i + 1] = rv[k][1] + beta * pv[k][1];
            }
        }
        for (int i = threadIdx.x + kCgReg * nt; i < P; i += nt) {
            p[2 * i] = r[2 * i] + beta * p[2 * i];
            p[2 * i + 1] = r[2 * i + 1] + beta * p[2 * i + 1];
        }
    }
    if (threadIdx.x == 0) {
        st->rs = sr;
        if (conv) {
            st->itn = st->it;
            st->done = 1;
        } else {
            st->rsold = sr;
            st->it += 1;
            if (st->it >= st->max_iter) {
                st->itn = st->max_iter;
                st->done = 1;
            }
        }
    }
}
__global__ void __launch_bounds__(1024) k_cg_update(const double* pdot, int nb, double* x, double* r, double* p,
                                                    const double* Ap, int P, CgStatus* st) {
    if (st->done) return;
    cg_update_body(pdot, nb, x, r, p, Ap, P, st);
}
// Ap = S p on the explicit S and the CG step in one launch: the block that
// finishes its rows last (device-scope release fence, then a counter; the
// counter returns to 0 for the next launch) runs the step.  Launches queued
// past convergence return at once.
__global__ void __launch_bounds__(256) k_dense_cg(const double* Sm, double* x, double* r, double* p, double* Ap, int P,
                                                  double* pdot, CgStatus* st, unsigned* cnt) {
    __shared__ int last;
    if (st->done) return;
    dense_mv_rows(Sm, p, Ap, P, pdot);
    if (threadIdx.x == 0) {
        __threadfence();  // this block's Ap rows and partial before the count
        last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every block's rows and partials are visible past the count
    cg_update_body(pdot, gridDim.x, x, r, p, Ap, P, st);
    if (threadIdx.x == 0) atomicExch(cnt, 0u);
}

// <O>, Var(O), <E O*> of the basic RBM from the structured sums:
// A = sum conj(O) and sum |t_i|^2 (mode 2), B = sum E conj(O) (mode 1),
// each as 32 chunk partials
__global__ void __launch_bounds__(256) k_struct_moments(const double* A, const double* B, int nch, int64_t S,
                                                        int N, int M, int P, double* mean, double* var,
                                                        double* eo, unsigned long long* maxvar) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int PA = P + M;
    auto sumA = [&](int q, double& re, double& im) {
        re = im = 0;
        for (int c = 0; c < nch; c++) {
            re += A[2 * ((int64_t)c * PA + q)];
            im += A[2 * ((int64_t)c * PA + q) + 1];
        }
    };
    double ar, ai, br = 0, bi = 0;
    sumA(p, ar, ai);
    for (int c = 0; c < nch; c++) {
        br += B[2 * ((int64_t)c * P + p)];
        bi += B[2 * ((int64_t)c * P + p) + 1];
    }
    double inv = 1.0 / (double)S;
    double mr = ar * inv, mi = -ai * inv;  // mean = conj(sum conj(O)) / S
    mean[2 * p] = mr;
    mean[2 * p + 1] = mi;
    eo[2 * p] = br * inv;
    eo[2 * p + 1] = bi * inv;
    if (p < N) {
        // sigma = +-1: sum_k |sigma_k - mean|^2 = S (1 - mean^2), where the
        // sum of sigma is an exact integer
        double v = S > 1 ? fmax((double)S * (1.0 - mr * mr) / ((double)S - 1.0), 0.0) : 0.0;
        var[p] = v;
        atomicMax(maxvar, (unsigned long long)__double_as_longlong(v));
    }
}

// Var(O) of the hidden and weight blocks of the structured basic RBM by a
// second pass over the tanh rows, once the means are known: sum_k |O_kp -
// mean_p|^2 with the deviation formed per sample (O = t_j for the hidden
// block, sigma_i t_j for the weights), so saturated units with |<O>|^2 >>
// Var(O) keep their digits.  Lane l = hidden unit 64 blockIdx.z + l;
// blockIdx.y = spins 16y .. 16y+15; the kVarWaves waves of a block walk chunks of kVarChunk
// samples in a grid-stride loop, the next chunk's loads issued before the
// current chunk is used.  FP32 deviations and chunk sums, FP64 across
// chunks; one partial row per block: [hidden j | W (i + j N)] (M + N M).
constexpr int kVarSpins = 16, kVarChunk = 8, kVarWaves = 4;
__global__ void __launch_bounds__(64 * kVarWaves) k_struct_var_part(const uint64_t* states, const float2* T,
                                                                    int64_t S, int N, int M, int nw,
                                                                    const double* mean, double* part) {
    __shared__ double red[kVarWaves - 1][kVarSpins + 1][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j0 = lane + 64 * (int)blockIdx.z;  // hidden units 64 z .. 64 z + 63
    const int j = j0 < M ? j0 : M - 1;           // lanes past M repeat unit M-1 and write nothing
    const int i0 = blockIdx.y * kVarSpins;
    const int ni = N - i0 < kVarSpins ? N - i0 : kVarSpins;
    float mr[kVarSpins], mi[kVarSpins];
    double a64[kVarSpins], h64 = 0;
#pragma unroll
    for (int q = 0; q < kVarSpins; q++) {
        const int p = N + M + (q < ni ? i0 + q : i0) + j * N;
        mr[q] = (float)mean[2 * p];
        mi[q] = (float)mean[2 * p + 1];
        a64[q] = 0;
    }
    const float hr = (float)mean[2 * (N + j)], hi = (float)mean[2 * (N + j) + 1];
    const int word = i0 >> 6, shift = i0 & 63;  // 16 spins never straddle a word
    const int64_t nch = (S + kVarChunk - 1) / kVarChunk, stride = (int64_t)gridDim.x * kVarWaves;
    float2 tc[kVarChunk];
    uint32_t bc[kVarChunk];
    auto fetch = [&](int64_t c, float2 (&t)[kVarChunk], uint32_t (&b)[kVarChunk]) {
#pragma unroll
        for (int r = 0; r < kVarChunk; r++) {
            const int64_t k0 = c * kVarChunk + r, k = k0 < S ? k0 : S - 1;  // clamped, masked below
            t[r] = T[k * M + j];
            b[r] = (uint32_t)(states[k * nw + word] >> shift);
        }
    };
    int64_t c = (int64_t)blockIdx.x * kVarWaves + w;
    if (c < nch) fetch(c, tc, bc);
    while (c < nch) {
        const int64_t cn = c + stride;
        float2 tn[kVarChunk];
        uint32_t bn[kVarChunk];
        if (cn < nch) fetch(cn, tn, bn);
        const int64_t nr = S - c * kVarChunk;
        float a32[kVarSpins], h32 = 0.f;
#pragma unroll
        for (int q = 0; q < kVarSpins; q++) a32[q] = 0.f;
#pragma unroll
        for (int r = 0; r < kVarChunk; r++) {
            if (r < nr) {
                const float2 t = tc[r];
#pragma unroll
                for (int q = 0; q < kVarSpins; q++) {
                    const float sg = (bc[r] >> q) & 1u ? 1.f : -1.f;  // -sigma
                    const float dr = fmaf(sg, mr[q], t.x), di = fmaf(sg, mi[q], t.y);
                    a32[q] = fmaf(dr, dr, fmaf(di, di, a32[q]));
                }
                const float dr = t.x - hr, di = t.y - hi;
                h32 = fmaf(dr, dr, fmaf(di, di, h32));
            }
        }
#pragma unroll
        for (int q = 0; q < kVarSpins; q++) a64[q] += a32[q];
        h64 += h32;
#pragma unroll
        for (int r = 0; r < kVarChunk; r++) {
            tc[r] = tn[r];
            bc[r] = bn[r];
        }
        c = cn;
    }
    // the block's waves add their sums in LDS
    if (w > 0) {
#pragma unroll
        for (int q = 0; q < kVarSpins; q++) red[w - 1][q][lane] = a64[q];
        red[w - 1][kVarSpins][lane] = h64;
    }
    __syncthreads();
    if (w == 0 && j0 < M) {
        for (int v = 0; v < kVarWaves - 1; v++) {
#pragma unroll
            for (int q = 0; q < kVarSpins; q++) a64[q] += red[v][q][lane];
            h64 += red[v][kVarSpins][lane];
        }
        double* row = part + (int64_t)blockIdx.x * (M + (int64_t)N * M);
        if (blockIdx.y == 0) row[j] = h64;
#pragma unroll
        for (int q = 0; q < kVarSpins; q++)
            if (q < ni) row[M + i0 + q + (int64_t)j * N] = a64[q];
    }
}

__global__ void __launch_bounds__(256) k_struct_var_final(const double* part, int64_t nchunk, int64_t S, int N,
                                                          int M, double* var, unsigned long long* maxvar) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, L = M + (int64_t)N * M;
    if (q >= L) return;
    double s = 0;
    for (int64_t c = 0; c < nchunk; c++) s += part[c * L + q];
    const double v = S > 1 ? s / ((double)S - 1.0) : 0.0;
    var[N + q] = v;
    atomicMax(maxvar, (unsigned long long)__double_as_longlong(v));
}

int ensure_sx(rbmgpu_ctx* c);
int run_struct(rbmgpu_ctx* c, int mode, const double* x, const CgStatus* st, hipEvent_t e0 = nullptr,
               hipEvent_t e1 = nullptr, void* part = nullptr);

int struct_moments(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (int rc = ensure_sx(c)) return rc;
    const int nch = rbmgpu_ctx::kSxChunks;
    double* A = c->d_sxpart2;
    double* B = A + (int64_t)nch * (c->P + c->na) * 2;
    SxArgs<float> a{};
    a.nblocks = c->sx_blocks;
    a.ld = (int)c->ld_s;
    a.nchunks = nch;
    a.part = (cplx<float>*)c->d_sxpart;
    // mode 2: sum conj(O) and sum |t|^2
    if (int rc = run_struct(c, 2, nullptr, nullptr)) return rc;
    a.P = (int)(c->P + c->na);
    a.part2 = A;
    HIPCK(launch_sx_reduce<float>(a, c->stream));
    // mode 1: sum E conj(O)
    if (int rc = run_struct(c, 1, nullptr, nullptr)) return rc;
    a.P = (int)c->P;
    a.part2 = B;
    HIPCK(launch_sx_reduce<float>(a, c->stream));
    if (int rc = zero_async(c, c->d_maxvar, sizeof(double))) return rc;
    hipLaunchKernelGGL(k_struct_moments, dim3((unsigned)((c->P + 255) / 256)), dim3(256), 0, c->stream, A, B, nch,
                       c->S, c->N, c->na, (int)c->P, c->d_omean, c->d_ovar, c->d_eo,
                       (unsigned long long*)c->d_maxvar);
    HIPCK(hipGetLastError());
    // second pass: Var(O) of the hidden and weight blocks
    const int64_t L = c->na + (int64_t)c->N * c->na;
    const int64_t nz = (c->na + 63) / 64;
    // about 3 waves per SIMD over the spin groups, at most one chunk per wave
    const int64_t ny = (c->N + kVarSpins - 1) / kVarSpins;
    const int64_t vch = (c->S + kVarChunk - 1) / kVarChunk;
    const int64_t nvb = std::max<int64_t>(
        1, std::min<int64_t>((12 * (int64_t)c->n_cu + ny * nz * kVarWaves - 1) / (ny * nz * kVarWaves),
                             (vch + kVarWaves - 1) / kVarWaves));
    const int64_t vneed = nvb * L * (int64_t)sizeof(double);
    if (vneed > c->varpart_cap) {
        if (int rc = dalloc(ctx, &c->d_varpart, vneed)) return rc;
        c->varpart_cap = vneed;
    }
    hipLaunchKernelGGL(k_struct_var_part, dim3((unsigned)nvb, (unsigned)ny, (unsigned)nz), dim3(64 * kVarWaves), 0,
                       c->stream,
                       c->d_states, (const float2*)c->d_T, c->S, c->N, c->na, c->nw, c->d_omean, c->d_varpart);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_struct_var_final, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, c->stream, c->d_varpart,
                       nvb, c->S, c->N, c->na, c->d_ovar, (unsigned long long*)c->d_maxvar);
    HIPCK(hipGetLastError());
    return 0;
}

int ensure_cg(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (c->d_x) return 0;
    size_t vb = c->P * 2 * sizeof(double);
    if (int rc = dalloc(ctx, &c->d_x, vb)) return rc;
    if (int rc = dalloc(ctx, &c->d_b, vb)) return rc;
    if (int rc = dalloc(ctx, &c->d_r, vb)) return rc;
    if (int rc = dalloc(ctx, &c->d_p, vb)) return rc;
    if (int rc = dalloc(ctx, &c->d_Ap, vb)) return rc;
    // partials of p^H Ap: one per S.x column block or per dense-S.x block of 4 rows
    const int64_t npd = std::max<int64_t>((c->P + kColBlock - 1) / kColBlock, (c->P + 3) / 4);
    if (int rc = dalloc(ctx, &c->d_pdot, npd * 2 * sizeof(double))) return rc;
    if (int rc = dalloc(ctx, &c->d_st, sizeof(CgStatus))) return rc;
    if (int rc = dalloc(ctx, &c->d_cgcnt, sizeof(unsigned))) return rc;
    HIPCK(hipMemsetAsync(c->d_cgcnt, 0, sizeof(unsigned), c->stream));  // not captured: allocation time
    return 0;
}

// S.x launch layout for the current sample set: block count and the byte
// size of the block partials (blocks x row complex)
int sx_layout(rbmgpu_ctx* c, int64_t* blocks_out, int64_t* bytes_out) {
    int64_t blocks, row;
    if (c->structured) {
        // chunks of 16 samples, one per wave step; the resident waves walk them
        int64_t chunks = (c->S + 15) / 16;
        int wpb = 4;
        int64_t bpc = std::max(struct_blocks_per_cu(c->N, c->na, &wpb), 1);
        blocks = std::min<int64_t>((chunks + wpb - 1) / wpb, bpc * c->n_cu);
        row = c->ld_s;
    } else {
        if (c->ld > sx_max_params()) return fail(c, "too many parameters for the S.x kernel");
        // 16 waves per CU (one block of 16, or four of 4 for long rows),
        // each wave streaming >= 8 rows
        const int wpb = sx_waves(c->ld);
        blocks = std::min<int64_t>(std::max<int64_t>(c->S / (8 * wpb), 1), (16 / wpb) * (int64_t)c->n_cu);
        row = c->ld;
    }
    *blocks_out = blocks;
    *bytes_out = blocks * row * 2 * (int64_t)rsize(c);
    return 0;
}

int ensure_sx(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    int64_t blocks, need;
    if (int rc = sx_layout(c, &blocks, &need)) return rc;
    c->sx_blocks = blocks;
    if (need > c->sxpart_cap) {
        if (int rc = dalloc(ctx, &c->d_sxpart, need)) return rc;
        c->sxpart_cap = need;
    }
    if (!c->d_sxpart2)
        if (int rc = dalloc(ctx, &c->d_sxpart2,
                            rbmgpu_ctx::kSxChunks * (2 * c->P + c->na) * 2 * sizeof(double)))
            return rc;
    return 0;
}

// argument block of the S.x reduction / finish (and of the materialised-O
// partial kernel)
template <typename R>
SxArgs<R> sx_args(rbmgpu_ctx* c, const double* x, double* y, double r1, double r2, double rd,
                  const CgStatus* st) {
    SxArgs<R> a{};
    a.reg_dev = c->reg_dev;
    a.wsx = c->weighted ? c->d_wsx : nullptr;
    a.S = c->S;
    a.nblocks = c->sx_blocks;
    a.P = (int)c->P;
    a.ld = (int)(c->structured ? c->ld_s : c->ld);
    a.nchunks = rbmgpu_ctx::kSxChunks;
    a.part2 = c->d_sxpart2;
    a.n_neural = (int)c->P_rbm;
    a.O = (const cplx<R>*)c->d_O;
    a.x = x;
    a.mean = c->d_omean;
    a.var = c->d_ovar;
    a.maxvar = c->d_maxvar;
    a.r1 = r1;
    a.r2 = r2;
    a.rd = rd;
    a.part = (cplx<R>*)c->d_sxpart;
    a.y = y;
    a.status = st;
    return a;
}

// block partials (the dominant kernel), reduction, finish, optional CG step
template <typename R>
int run_sx(rbmgpu_ctx* c, const double* x, double* y, double r1, double r2, double rd, const CgStatus* st,
           double* pdot, bool cg, void* part) {
    rbmgpu_ctx* ctx = c;
    SxArgs<R> a = sx_args<R>(c, x, y, r1, r2, rd, st);
    if (part) a.part = (cplx<R>*)part;
    if (c->structured) {
        if (int rc = run_struct(c, 0, x, st, nullptr, nullptr, part)) return rc;
    } else {
        HIPCK(launch_sx<R>(a, (int)c->sx_blocks, c->stream));
    }
    a.pdot = pdot;
    HIPCK(launch_sx_colsum<R>(a, c->stream));
    if (cg) {
        hipLaunchKernelGGL(k_cg_update, dim3(1), dim3(1024), 0, c->stream, pdot,
                           (int)((c->P + kColBlock - 1) / kColBlock), c->d_x, c->d_r, c->d_p, c->d_Ap, (int)c->P,
                           c->d_st);
        HIPCK(hipGetLastError());
    }
    return 0;
}

int run_struct(rbmgpu_ctx* c, int mode, const double* x, const CgStatus* st, hipEvent_t e0, hipEvent_t e1,
               void* part) {
    rbmgpu_ctx* ctx = c;
    StructArgs a{};
    a.S = c->S;
    a.N = c->N;
    a.M = c->na;
    a.P = (int)c->P;
    a.ld = (int)c->ld_s;
    a.nw = c->nw;
    a.n_groups = c->n_groups;
    a.mode = mode;
    a.states = c->d_states;
    a.T = (const cplx<float>*)c->d_T;
    a.obs = c->d_obs;
    a.x = x;
    a.part = (cplx<float>*)(part ? part : c->d_sxpart);
    a.status = st;
    HIPCK(launch_struct(a, (int)c->sx_blocks, c->stream, e0, e1));
    return 0;
}

// pdot: optional per-column-block partials of x^H y for the CG update
// cg: the CG step on the product follows (x = d_p, y = d_Ap, pdot given)
// part: block-partial buffer (default: the context's own)
int sx(rbmgpu_ctx* c, const double* x, double* y, double r1, double r2, double rd, const CgStatus* st,
       double* pdot = nullptr, bool cg = false, void* part = nullptr) {
    return c->prec == RBMGPU_FP32 ? run_sx<float>(c, x, y, r1, r2, rd, st, pdot, cg, part)
                                  : run_sx<double>(c, x, y, r1, r2, rd, st, pdot, cg, part);
}

int compute_moments(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (c->S == 0) return fail(c, "no samples");
    if (!c->d_emom) {
        if (int rc = dalloc(ctx, &c->d_emom, (16 + 2 * c->n_groups) * sizeof(double))) return rc;
        if (int rc = dalloc(ctx, &c->d_omean, c->P * 2 * sizeof(double))) return rc;
        if (int rc = dalloc(ctx, &c->d_ovar, c->P * sizeof(double))) return rc;
        if (int rc = dalloc(ctx, &c->d_eo, c->P * 2 * sizeof(double))) return rc;
        if (int rc = dalloc(ctx, &c->d_maxvar, sizeof(double))) return rc;
    }
    // bins: the largest count <= 50 that divides S (abstract_optimizer.cpp:188-190);
    // ranges of about 1024 samples that tile every bin
    int64_t nb = 50;
    while (c->S % nb != 0) nb--;
    const int64_t unit = nb > 1 ? c->S / nb : c->S;
    int64_t chunk = 1024;
    if (nb > 1) {
        chunk = unit;
        for (int64_t d = std::min<int64_t>(unit, 1024); d >= 64; d--)
            if (unit % d == 0) {
                chunk = d;
                break;
            }
    }
    const int64_t per_bin = nb > 1 ? unit / chunk : 0;
    const int64_t nblk = nb > 1 ? nb * per_bin : (c->S + chunk - 1) / chunk;
    const int64_t pneed = nblk * (3 + 2 * c->n_groups) * (int64_t)sizeof(double);
    if (pneed > c->epart_cap) {
        if (int rc = dalloc(ctx, &c->d_epart, pneed)) return rc;
        c->epart_cap = pneed;
    }
    if (int rc = record(c, 4)) return rc;
    const double* wts = c->weighted ? c->d_w : nullptr;
    hipLaunchKernelGGL(k_e_part, dim3((unsigned)nblk), dim3(256), 0, c->stream, c->d_obs, wts, c->S, c->n_groups,
                       chunk, c->d_epart);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_e_final, dim3(1), dim3(1024), 0, c->stream, c->d_epart, nblk, per_bin, c->S, c->n_groups,
                       nb, c->d_emom);
    HIPCK(hipGetLastError());
    if (c->structured) {
        if (int rc = struct_moments(c)) return rc;
    } else {
        int64_t chunk = 256;
        int64_t blocks = (c->S + chunk - 1) / chunk;
        if (blocks > 4096) {
            blocks = 4096;
            chunk = (c->S + blocks - 1) / blocks;
            blocks = (c->S + chunk - 1) / chunk;
        }
        int64_t need = blocks * c->P * kOPartRow;
        if (need > c->part_cap) {
            if (int rc = dalloc(ctx, &c->d_part, need * sizeof(double))) return rc;
            c->part_cap = need;
        }
        if (c->prec == RBMGPU_FP32) {
            MomArgs<float> a{c->S, chunk, (int)c->P, c->n_groups, (int)c->ld, (const cplx<float>*)c->d_O, c->d_obs,
                             c->d_part, wts};
            HIPCK(launch_o_partial<float>(a, (int)blocks, c->stream));
        } else {
            MomArgs<double> a{c->S, chunk, (int)c->P, c->n_groups, (int)c->ld, (const cplx<double>*)c->d_O, c->d_obs,
                              c->d_part, wts};
            HIPCK(launch_o_partial<double>(a, (int)blocks, c->stream));
        }
        if (int rc = zero_async(c, c->d_maxvar, sizeof(double))) return rc;
        hipLaunchKernelGGL(k_o_final, dim3((unsigned)c->P), dim3(256), 0, c->stream, c->d_part, blocks, (int)c->P,
                           c->d_omean, c->d_ovar, c->d_eo, (unsigned long long*)c->d_maxvar);
        HIPCK(hipGetLastError());
    }
    if (int rc = record(c, 5)) return rc;
    c->moments_valid = true;
    return 0;
}

int ensure_eval(rbmgpu_ctx* c, int64_t n) {
    rbmgpu_ctx* ctx = c;
    if (n <= c->ev_cap) return 0;
    if (int rc = dalloc(ctx, &c->d_evst, n * c->nw * sizeof(uint64_t))) return rc;
    if (int rc = dalloc(ctx, &c->d_evout, n * 2 * sizeof(double))) return rc;
    c->ev_cap = n;
    return 0;
}

template <typename R>
int run_eval(rbmgpu_ctx* c, const uint64_t* d_states, int64_t n, double* logpsi, double* obs, void* Tout,
             double* ratio, const int32_t* flips, int nf) {
    rbmgpu_ctx* ctx = c;
    EvalArgs<R> a{};
    a.t = tabs<R>(c);
    a.T = terms(c);
    a.n = n;
    a.states = d_states;
    a.logpsi = logpsi;
    a.obs = obs;
    a.Tout = (cplx<R>*)Tout;
    a.ratio = ratio;
    // the sample-set path (T written into the context) also writes the
    // Pfaffian derivatives next to it
    a.Opf = (c->pf_rows && Tout == c->d_T) ? (cplx<R>*)c->d_Opf : nullptr;
    a.pb = c->d_pb;
    a.n_flips = nf;
    for (int i = 0; i < nf; i++) a.flips[i] = flips[i];
    Shape s = c->shape.L && !c->shape.wave ? c->shape : Shape{0, 0};
    if (!s.L) {
        for (int i = 0; i < kNumShapes; i++)
            if (kShapes[i].L == 64 && shape_fits(c, kShapes[i]) && (!s.L || kShapes[i].K < s.K)) s = kShapes[i];
        if (!s.L)
            for (int i = 0; i < kNumShapes; i++)
                if (shape_fits(c, kShapes[i]) && (!s.L || kShapes[i].L * kShapes[i].K < s.L * s.K)) s = kShapes[i];
    }
    if (!s.L) return fail(c, "no launch shape for this number of hidden units");
    int block = pick_block(c, s.L, n);
    a.d = dims(c, block);
    HIPCK(launch_eval<R>(a, s, block, c->stream));
    return 0;
}

int eval(rbmgpu_ctx* c, const uint64_t* d_states, int64_t n, double* logpsi, double* obs, void* Tout,
         double* ratio, const int32_t* flips, int nf) {
    return c->prec == RBMGPU_FP32 ? run_eval<float>(c, d_states, n, logpsi, obs, Tout, ratio, flips, nf)
                                  : run_eval<double>(c, d_states, n, logpsi, obs, Tout, ratio, flips, nf);
}

// ---- device-resident SR epoch (rbmgpu_train_epoch): sample -> moments ->
// F (SR.cpp:102) -> CG (cg_solver.cpp:119-189) -> massive-step guard and
// w -= lr x (abstract_optimizer.cpp:153-172, rbm_base.cpp:117-143) ->
// weight tables, with every per-epoch value (epoch, lr, regularisation,
// iteration cap) read from DevPar, so the whole body can be captured once
// in a HIP graph and replayed.

// F = <E O*> - <E> conj(<O>)  (SR.cpp:102)
__global__ void __launch_bounds__(256) k_sr_grad(const double* emom, const double* eo, const double* om, int P,
                                                 double* b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    {
#pragma clang fp contract(off)  // each product rounded on its own, as the host computes F
        const double er = emom[0], ei = emom[1], mr = om[2 * p], mi = om[2 * p + 1];
        b[2 * p] = eo[2 * p] - (er * mr + ei * mi);  // E conj(mu) = (er + i ei)(mr - i mi)
        b[2 * p + 1] = eo[2 * p + 1] - (ei * mr - er * mi);
    }
}

// after a finished CG solve: the guard max |x| <= 30 with every entry
// finite, then x <- lr x (the next epoch's warm start, as the reference
// scales dw_ in place) and w -= x.  One block.  A solve that has not
// finished leaves everything as it is and flags `pending`.
__global__ void __launch_bounds__(1024) k_sr_update(const DevPar* par, const CgStatus* st, double* x,
                                                    double* params, int P, DevStats* out) {
    __shared__ double smx[16];
    __shared__ int sbad[16];
    if (!st->done) {
        if (threadIdx.x == 0) {
            out->pending = 1;
            out->ok = 0;
        }
        return;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double mx = 0;
    int bad = 0;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const double a = hypot(x[2 * i], x[2 * i + 1]);
        if (!isfinite(a)) bad = 1;
        else mx = fmax(mx, a);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        bad |= __shfl_xor(bad, o);
    }
    if (lane == 0) {
        smx[w] = mx;
        sbad[w] = bad;
    }
    __syncthreads();
    mx = 0;
    bad = 0;
    for (int q = 0; q < nw; q++) {
        mx = fmax(mx, smx[q]);
        bad |= sbad[q];
    }
    const bool ok = !bad && mx <= 30.0;
    if (ok) {
#pragma clang fp contract(off)  // w - (lr x) with lr x rounded first, as on the host
        const double lr = par->lr;
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
            const double xr = lr * x[2 * i], xi = lr * x[2 * i + 1];
            x[2 * i] = xr;
            x[2 * i + 1] = xi;
            params[2 * i] -= xr;
            params[2 * i + 1] -= xi;
        }
    }
    if (threadIdx.x == 0) {
        out->ok = ok;
        out->pending = 0;
        out->max_dw = bad ? (double)NAN : mx;
        out->itn = st->itn;
        out->rs = st->rs;
    }
}

// weight tables from the device parameters (the layout of
// build_weight_tables), after an accepted update only
template <typename R>
__global__ void __launch_bounds__(256) k_tables(const double* params, int N, int na, int nvb, int nsymm,
                                                const int32_t* vbsite, int pf_rows, int64_t P_rbm,
                                                const DevStats* st, cplx<R>* W, cplx<R>* C, cplx<R>* Sh,
                                                cplx<R>* hb, cplx<R>* rb, cplx<R>* lb, cplx<R>* fs) {
    if (!st->ok) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nwt = (int64_t)N * na;
    if (t < nwt) {
        const int k = (int)(t / na), a = (int)(t % na);
        const int64_t q = nvb + na + k + (int64_t)a * N;
        const double x = 2 * params[2 * q], y = 2 * params[2 * q + 1];
        double sy, cy;
        sincos(y, &sy, &cy);
        const double chx = cosh(x), shx = sinh(x);
        W[t] = {(R)params[2 * q], (R)params[2 * q + 1]};
        C[t] = {(R)(chx * cy), (R)(shx * sy)};   // cosh(x + iy)
        Sh[t] = {(R)(shx * cy), (R)(chx * sy)};  // sinh(x + iy)
    } else if (t < nwt + na) {
        const int a = (int)(t - nwt);
        hb[a] = {(R)params[2 * (nvb + a)], (R)params[2 * (nvb + a) + 1]};
    } else if (t < nwt + na + N) {
        const int f = (int)(t - nwt - na);
        rb[f] = {(R)params[2 * (f % nvb)], (R)params[2 * (f % nvb) + 1]};
        double lr = 0, li = 0;  // lb[f]: sum of v_i over the symmetries that map basis i to f
        for (int s = 0; s < nsymm; s++)
            for (int i = 0; i < nvb; i++)
                if (vbsite[s * nvb + i] == f) {
                    lr += params[2 * i];
                    li += params[2 * i + 1];
                }
        lb[f] = {(R)lr, (R)li};
    } else if (t < nwt + na + N + 4 * (int64_t)pf_rows) {
        const int e = (int)(t - nwt - na - N), r = e / 4, q = e % 4;
        const int64_t pi = P_rbm + r + (int64_t)q * pf_rows;
        fs[e] = {(R)params[2 * pi], (R)params[2 * pi + 1]};
    }
}

// energy statistics and the mean chain acceptance of the epoch
__global__ void __launch_bounds__(1024) k_epoch_stats(const double* emom, const uint32_t* acc,
                                                      const uint32_t* prop, int n_chains, DevStats* out) {
    double a = 0, dummy = 0;
    for (int i = threadIdx.x; i < n_chains; i += blockDim.x) a += prop[i] ? (double)acc[i] / prop[i] : 0.0;
    block_sum2(a, dummy);
    if (threadIdx.x == 0) {
        for (int q = 0; q < 6; q++) out->e[q] = emom[q];
        out->acc = n_chains > 0 ? a / n_chains : 0.0;
    }
}

// explicit S of the current sample set (kern_dense.hip) into d_Smat: mode 0
// with the CG regularisation (cg_solver.cpp:276-281), mode 1 with the
// direct solver's (direct_solver.cpp:45-53)
template <typename R>
int build_S_t(rbmgpu_ctx* c, double r1, double r2, double rd, int mode, const double* reg_dev) {
    rbmgpu_ctx* ctx = c;
    SyrkArgs<R> a{};
    a.X = (const R*)c->d_O;
    a.S = c->S;
    a.ncol = (int)(2 * c->ld);
    a.nvalid = (int)(2 * c->P);
    a.ntile = (int)((2 * c->P + 63) / 64);
    a.npairs = a.ntile * (a.ntile + 1) / 2;
    // about 7 blocks per CU over tiles x sample slices of >= 256 samples
    // (4 stay resident per CU; dynamic dispatch then evens out the CUs),
    // the slice count a multiple of 8 where the samples allow, so every
    // slice stays on one XCD
    int64_t ks = std::max<int64_t>(1, (7 * (int64_t)c->n_cu + a.npairs - 1) / a.npairs);
    ks = std::min<int64_t>(ks, std::max<int64_t>(1, c->S / 256));
    if (ks >= 8) ks = ks / 8 * 8;
    a.kper = ((c->S + ks - 1) / ks + 31) / 32 * 32;
    a.ksplit = (int)((c->S + a.kper - 1) / a.kper);
    if (a.ksplit % 8 && ks % 8 == 0) a.ksplit = (int)ks;  // trailing slices may be empty
    const int64_t gbytes = (int64_t)a.npairs * 64 * 64 * (int64_t)sizeof(double);
    const int64_t need = gbytes + (int64_t)a.ksplit * a.npairs * 64 * 64 * (int64_t)sizeof(R);
    if (need > c->gpart_cap) {
        if (int rc = dalloc(ctx, &c->d_gpart, need)) return rc;
        c->gpart_cap = need;
    }
    a.G = (double*)c->d_gpart;
    a.part = (R*)((char*)c->d_gpart + gbytes);
    a.wsx = c->weighted ? c->d_wsx : nullptr;
    a.mean = c->d_omean;
    HIPCK(launch_syrk<R>(a, c->stream));
    HIPCK(launch_s_from_g<R>(a, (int)c->P, (int)c->P_rbm, c->S, c->d_omean, c->d_ovar, c->d_maxvar, r1, r2, rd,
                             reg_dev, mode, c->d_Smat, c->stream));
    return 0;
}

int build_S(rbmgpu_ctx* c, double r1, double r2, double rd, int mode, const double* reg_dev = nullptr) {
    rbmgpu_ctx* ctx = c;
    if (c->structured || !c->d_O)
        return fail(c, "explicit S needs the O matrix: set keep_o (gpu.keep_o = true) for the cg-single / direct "
                       "solvers");
    if (c->S < 1) return fail(c, "no samples");
    if (!c->moments_valid)
        if (int rc = compute_moments(c)) return rc;
    if (!c->d_Smat)
        if (int rc = dalloc(ctx, &c->d_Smat, c->P * c->P * 16)) return rc;
    return c->prec == RBMGPU_FP32 ? build_S_t<float>(c, r1, r2, rd, mode, reg_dev)
                                  : build_S_t<double>(c, r1, r2, rd, mode, reg_dev);
}

// x = A^-1 b for the dense Hermitian A (column-major) on the host: Cholesky
// A = L L^H, or LU with partial pivoting when A is not positive definite.
// The direct solver's regularised S is positive definite for r2 > 0, where
// the reference's complete orthogonal decomposition gives the same x.
int dense_solve(rbmgpu_ctx* c, const std::vector<zd>& A, const std::vector<zd>& b, std::vector<zd>& x) {
    const int64_t n = (int64_t)b.size();
    std::vector<zd> L(A);
    bool pd = true;
    for (int64_t j = 0; j < n && pd; j++) {
        double d = L[j + j * n].real();
        for (int64_t k = 0; k < j; k++) d -= std::norm(L[j + k * n]);
        if (!(d > 0)) {
            pd = false;
            break;
        }
        d = std::sqrt(d);
        L[j + j * n] = d;
        for (int64_t i = j + 1; i < n; i++) {
            zd s = L[i + j * n];
            for (int64_t k = 0; k < j; k++) s -= L[i + k * n] * std::conj(L[j + k * n]);
            L[i + j * n] = s / d;
        }
    }
    if (pd) {
        std::vector<zd> y(n);
        for (int64_t i = 0; i < n; i++) {
            zd s = b[i];
            for (int64_t k = 0; k < i; k++) s -= L[i + k * n] * y[k];
            y[i] = s / L[i + i * n].real();
        }
        for (int64_t i = n - 1; i >= 0; i--) {
            zd s = y[i];
            for (int64_t k = i + 1; k < n; k++) s -= std::conj(L[k + i * n]) * x[k];
            x[i] = s / L[i + i * n].real();
        }
        return 0;
    }
    std::vector<zd> M(A), r(b);
    std::vector<int64_t> piv(n);
    for (int64_t j = 0; j < n; j++) {
        int64_t p = j;
        for (int64_t i = j + 1; i < n; i++)
            if (std::abs(M[i + j * n]) > std::abs(M[p + j * n])) p = i;
        if (std::abs(M[p + j * n]) == 0) return fail(c, "direct solver: singular S");
        if (p != j) {
            for (int64_t k = 0; k < n; k++) std::swap(M[j + k * n], M[p + k * n]);
            std::swap(r[j], r[p]);
        }
        for (int64_t i = j + 1; i < n; i++) {
            const zd f = M[i + j * n] / M[j + j * n];
            if (f == zd(0)) continue;
            for (int64_t k = j; k < n; k++) M[i + k * n] -= f * M[j + k * n];
            r[i] -= f * r[j];
        }
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        zd s = r[i];
        for (int64_t k = i + 1; k < n; k++) s -= M[i + k * n] * x[k];
        x[i] = s / M[i + i * n];
    }
    return 0;
}

// ---- exact enumeration (full_sampler.cpp:43-185): all 2^N states in the
// reference's order (2^bp Gray-code runs over the upper N - bp bits, run b
// starting from the state b), each weighted by |psi|^2
__global__ void __launch_bounds__(256) k_gray_states(uint64_t* states, int64_t S, int nw, int bp, int N) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    const int64_t run = k >> (N - bp), i = k & ((((int64_t)1) << (N - bp)) - 1);
    const uint64_t bits = (uint64_t)run | ((uint64_t)(i ^ (i >> 1)) << bp);
    states[k * nw] = bits;
    for (int q = 1; q < nw; q++) states[k * nw + q] = 0;
}

// per-block max of Re log psi
__global__ void __launch_bounds__(256) k_lp_max(const double* lp, int64_t S, double* part) {
    __shared__ double red[4];
    double m = -INFINITY;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S; k += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, lp[2 * k]);
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// w_k = exp(2 (Re log psi_k - max)) = |psi_k|^2 / max |psi|^2 and the block sums
__global__ void __launch_bounds__(256) k_weights(const double* lp, int64_t S, const double* pmax, int nb, double* w,
                                                 double* psum) {
    double m = -INFINITY;
    for (int b = 0; b < nb; b++) m = fmax(m, pmax[b]);
    double s = 0, dummy = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S; k += (int64_t)gridDim.x * blockDim.x) {
        const double v = exp(2.0 * (lp[2 * k] - m));
        w[k] = v;
        s += v;
    }
    block_sum2(s, dummy);
    if (threadIdx.x == 0) psum[blockIdx.x] = s;
}

// S w_k / sum w: the per-sample factor of the weighted S.x (the O matrix of
// outer_aggregator_lazy holds sqrt(w) O, aggregator.cpp:270-305)
__global__ void __launch_bounds__(256) k_wsx(const double* w, int64_t S, const double* psum, int nb, double* wsx) {
    double W = 0;
    for (int b = 0; b < nb; b++) W += psum[b];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S; k += (int64_t)gridDim.x * blockDim.x)
        wsx[k] = (double)S * w[k] / W;
}

// one sampling epoch on the stream: all chains, the kept samples' local
// observables and (materialised path) the O matrix; buffers allocated
int sample_body(rbmgpu_ctx* c, uint32_t epoch) {
    rbmgpu_ctx* ctx = c;
    c->weighted = false;
    if (int rc = zero_async(c, c->d_obs, c->n_samples * 2 * c->n_groups * sizeof(double))) return rc;
    if (int rc = record(c, 0)) return rc;
    int rc = 0;
    if (c->Hn && c->lam > 0) {
        LearnArgs la = learn_args(c, epoch);
        la.s.epoch_dev = c->epoch_dev;
        HIPCK(launch_sample_mlp(la, learn_k(c), c->stream));
    } else {
        rc = c->prec == RBMGPU_FP32 ? run_sample<float>(c, epoch, c->shape, c->block, c->n_samples, c->warmup)
                                    : run_sample<double>(c, epoch, c->shape, c->block, c->n_samples, c->warmup);
        if (!rc && eloc_after(c)) rc = run_eloc(c, c->n_samples);
    }
    if (rc) return rc;
    if (int rc2 = record(c, 1)) return rc2;
    c->learned_epoch = c->Hn && c->lam > 0;
    c->S = c->n_samples;
    c->have_chains = true;
    rc = c->prec == RBMGPU_FP32 ? run_build_O<float>(c) : run_build_O<double>(c);
    if (rc) return rc;
    if (int rc2 = record(c, 2)) return rc2;
    c->moments_valid = false;
    return 0;
}


int ensure_train(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (!c->d_par) {
        if (int rc = dalloc(ctx, &c->d_par, sizeof(DevPar))) return rc;
        if (int rc = dalloc(ctx, &c->d_stats, sizeof(DevStats))) return rc;
        HIPCK(hipHostMalloc(&c->h_par, sizeof(DevPar), hipHostMallocDefault));
        HIPCK(hipHostMalloc(&c->h_stats, sizeof(DevStats), hipHostMallocDefault));
        // the reference's dw_ starts at zero (SR.cpp:57)
        if (int rc = zero_async(c, c->d_x, c->P * 2 * sizeof(double))) return rc;
    }
    return 0;
}

// guard, update, tables and statistics after the CG launches
int train_tail(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    hipLaunchKernelGGL(k_sr_update, dim3(1), dim3(1024), 0, c->stream, (const DevPar*)c->d_par, c->d_st, c->d_x,
                       c->d_params, (int)c->P, (DevStats*)c->d_stats);
    HIPCK(hipGetLastError());
    const int64_t nt = (int64_t)c->N * c->na + c->na + c->N + 4 * (int64_t)c->pf_rows;
    if (c->prec == RBMGPU_FP32) {
        Tabs<float> t = tabs<float>(c);
        hipLaunchKernelGGL(k_tables<float>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, c->d_params,
                           c->N, c->na, c->nvb, c->nsymm, c->d_vbsite, c->pf_rows, c->P_rbm,
                           (const DevStats*)c->d_stats, (cplx<float>*)t.W, (cplx<float>*)t.C, (cplx<float>*)t.S,
                           (cplx<float>*)t.hb, (cplx<float>*)t.rb, (cplx<float>*)t.lb, (cplx<float>*)t.fs);
    } else {
        Tabs<double> t = tabs<double>(c);
        hipLaunchKernelGGL(k_tables<double>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_params, c->N, c->na, c->nvb, c->nsymm, c->d_vbsite, c->pf_rows, c->P_rbm,
                           (const DevStats*)c->d_stats, (cplx<double>*)t.W, (cplx<double>*)t.C,
                           (cplx<double>*)t.S, (cplx<double>*)t.hb, (cplx<double>*)t.rb, (cplx<double>*)t.lb,
                           (cplx<double>*)t.fs);
    }
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_epoch_stats, dim3(1), dim3(1024), 0, c->stream, c->d_emom, c->d_acc, c->d_prop, c->n_chains,
                       (DevStats*)c->d_stats);
    HIPCK(hipGetLastError());
    return 0;
}

// the per-epoch values come from DevPar while an epoch is enqueued
struct DevArgsScope {
    rbmgpu_ctx* c;
    explicit DevArgsScope(rbmgpu_ctx* cc) : c(cc) {
        c->epoch_dev = (const uint32_t*)((const char*)c->d_par + offsetof(DevPar, epoch));
        c->reg_dev = (const double*)((const char*)c->d_par + offsetof(DevPar, reg));
    }
    ~DevArgsScope() {
        c->epoch_dev = nullptr;
        c->reg_dev = nullptr;
    }
};

// the whole epoch with n_cg CG iterations queued (launches past convergence
// return at once)
// one matrix-vector product of the SR epoch's CG: the regularised S.x
// matrix-free, or the explicit S (train_dense); cg: the CG step follows
// Ap = S p and the CG step on the explicit S, one fused launch
int dense_cg_step(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    hipLaunchKernelGGL(k_dense_cg, dim3((unsigned)((c->P + 3) / 4)), dim3(256), 0, c->stream, c->d_Smat, c->d_x,
                       c->d_r, c->d_p, c->d_Ap, (int)c->P, c->d_pdot, c->d_st, c->d_cgcnt);
    HIPCK(hipGetLastError());
    return 0;
}

int train_mv(rbmgpu_ctx* c, const double* x, double* y, const CgStatus* st, bool cg) {
    rbmgpu_ctx* ctx = c;
    if (c->train_dense != 1) return sx(c, x, y, 0, 0, 0, st, cg ? c->d_pdot : nullptr, cg);
    if (cg) return dense_cg_step(c);
    HIPCK(launch_dense_mv(c->d_Smat, x, y, (int)c->P, nullptr, st, c->stream));
    return 0;
}

int train_body(rbmgpu_ctx* c, int64_t n_cg) {
    rbmgpu_ctx* ctx = c;
    DevArgsScope scope(c);
    if (int rc = sample_body(c, 0)) return rc;
    if (int rc = ensure_sx(c)) return rc;  // S.x layout of this sample set (allocates only when it grows)
    if (int rc = compute_moments(c)) return rc;
    hipLaunchKernelGGL(k_sr_grad, dim3((unsigned)((c->P + 255) / 256)), dim3(256), 0, c->stream, c->d_emom, c->d_eo,
                       c->d_omean, (int)c->P, c->d_b);
    HIPCK(hipGetLastError());
    if (c->train_dense == 1) {
        // the explicit S with this epoch's regularisation (DevPar::reg)
        const double* reg = reinterpret_cast<const double*>(reinterpret_cast<const char*>(c->d_par) +
                                                            offsetof(DevPar, reg));
        if (int rc = build_S(c, 0, 0, 0, 0, reg)) return rc;
    }
    if (int rc = train_mv(c, c->d_x, c->d_Ap, nullptr, false)) return rc;
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1024), 0, c->stream, c->d_b, c->d_Ap, c->d_r, c->d_p, (int)c->P,
                       c->d_st, (int64_t)0, 0.0, (const DevPar*)c->d_par);
    HIPCK(hipGetLastError());
    for (int64_t i = 0; i < n_cg; i++)
        if (int rc = train_mv(c, c->d_p, c->d_Ap, c->d_st, true)) return rc;
    return train_tail(c);
}

// more CG iterations of an unfinished solve, then the tail again
int train_more(rbmgpu_ctx* c, int64_t n_cg) {
    DevArgsScope scope(c);
    for (int64_t i = 0; i < n_cg; i++)
        if (int rc = train_mv(c, c->d_p, c->d_Ap, c->d_st, true)) return rc;
    return train_tail(c);
}

// the SR epoch's CG on the explicit S when that is cheaper: one SYRK of O
// (2 S (2P)^2 flops: MFMA at ~100 TFLOP/s of that count in FP32 -- the
// upper-triangle tiles do about 2/3 of it --, VALU at ~15 in FP64) and then
// per iteration a P x P complex mat-vec fused with the CG step (~8 us)
// against a pass over O per iteration (S ld 8 or 16 bytes at ~6 TB/s plus
// ~15 us of launches); the same trade as the reference's cg-dynamic
// (cg_solver.cpp:31-46), with this device's rates
int decide_train_dense(const rbmgpu_ctx* c, int64_t max_iter) {
    if (c->structured) return 0;  // no O matrix
    const double S = (double)c->n_samples, P = (double)c->P, ld = (double)c->ld;
    const double rsz = c->prec == RBMGPU_FP32 ? 8.0 : 16.0;
    const double syrk = 2.0 * S * (2 * P) * (2 * P) / (c->prec == RBMGPU_FP32 ? 100e12 : 15e12);
    const double dense = syrk + (double)max_iter * (P * P * 16.0 / 4e12 + 8e-6);
    const double mfree = (double)max_iter * (S * ld * rsz / 6e12 + 15e-6);
    return dense < mfree ? 1 : 0;
}

void drop_graphs(rbmgpu_ctx* c) {
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
}

}  // namespace

// ====================================================================== ABI
extern "C" {

const char* rbmgpu_last_error(const rbmgpu_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rbmgpu_create(const rbmgpu_config* cfg, rbmgpu_ctx** out) {
    if (!cfg || !out) return -1;
    *out = nullptr;
    rbmgpu_ctx* ctx = new rbmgpu_ctx();
    rbmgpu_ctx* c = ctx;
    auto bad = [&](const std::string& m) {
        *out = c;  // keep the context so the caller can read the message
        return fail(c, m);
    };
    c->N = cfg->n_spins;
    c->nvb = cfg->n_vb;
    c->na = cfg->n_alpha;
    c->nsymm = cfg->n_symm;
    if (c->N < 1 || c->N > 128) return bad("n_spins must be in 1..128 (MAX_SPIN_SITES)");
    if (c->nvb < 1 || c->na < 1 || c->nsymm < 1) return bad("n_vb, n_alpha and n_symm must be positive");
    if (c->nsymm > 1 && c->N > 255) return bad("permuted sites must fit a byte");
    c->H = c->na * c->nsymm;
    c->nw = (c->N + 63) / 64;
    c->P_rbm = (int64_t)c->nvb + c->na + (int64_t)c->na * c->N;
    c->pf_rows = cfg->pf_rows;
    c->pf_refresh = cfg->pf_refresh;
    std::vector<int16_t> bs;
    if (c->pf_rows < 0) return bad("pf_rows must be >= 0");
    if (c->pf_rows) {
        if (c->N > kPfMaxN) return bad("the Pfaffian factor supports at most 32 spins");
        if (!cfg->pf_bs) return bad("pf_bs is required with a Pfaffian factor");
        bs.resize((size_t)c->N * c->N);
        for (int i = 0; i < c->N; i++)
            for (int j = 0; j < c->N; j++) {
                int v = cfg->pf_bs[i + (int64_t)j * c->N];  // column-major b(i, j)
                if ((i == j) != (v == 0) || std::abs(v) > c->pf_rows) return bad("invalid Pfaffian index map");
                bs[(size_t)i * c->N + j] = (int16_t)v;
            }
    }
    c->P = c->P_rbm + 4 * (int64_t)c->pf_rows;
    c->ld = c->P + (c->P & 1);
    if (!cfg->perm || !cfg->basis) return bad("perm and basis tables are required");
    c->perm.assign(cfg->perm, cfg->perm + (int64_t)c->nsymm * c->N);
    c->basis.assign(cfg->basis, cfg->basis + c->nvb);
    for (int s = 0; s < c->nsymm; s++) {
        std::vector<int> seen(c->N, 0);
        for (int f = 0; f < c->N; f++) {
            int v = c->perm[s * c->N + f];
            if (v < 0 || v >= c->N || seen[v]++) return bad("perm row is not a permutation");
        }
    }
    for (int b : c->basis)
        if (b < 0 || b >= c->N) return bad("basis site out of range");
    // ---- sampler settings
    c->n_chains = cfg->n_chains;
    if (c->n_chains == 0 && cfg->n_samples > 0) {
        // automatic chain count: start from min(n_samples, 16384), candidates
        // up to min(n_samples, 65536)
        c->auto_chains = true;
        c->n_chains = (int)std::min<int64_t>(cfg->n_samples, 16384);
    }
    c->chains_cap = c->auto_chains ? (int)std::min<int64_t>(cfg->n_samples, 65536) : c->n_chains;
    c->n_samples = cfg->n_samples;
    c->warmup = cfg->n_warmup;
    c->step = cfg->step_size;
    c->bond_flips = cfg->bond_flips;
    c->seed = cfg->seed;
    c->prec = cfg->precision;
    c->lanes_req = cfg->lanes_per_chain;
    c->train_dense = cfg->sr_explicit_s == 0 || cfg->sr_explicit_s == 1 ? cfg->sr_explicit_s : -1;
    c->keep_o = cfg->keep_o;
    if (c->prec != RBMGPU_FP32 && c->prec != RBMGPU_FP64) return bad("precision must be 32 or 64");
    if (c->n_chains < 1 || c->n_samples < 1 || c->step < 1 || c->warmup < 0)
        return bad("invalid sampler settings");
    if (cfg->n_prop_bonds > 0) {
        if (!cfg->prop_bonds) return bad("prop_bonds missing");
        c->pb.assign(cfg->prop_bonds, cfg->prop_bonds + 2 * cfg->n_prop_bonds);
        for (int v : c->pb)
            if (v < 0 || v >= c->N) return bad("proposal bond site out of range");
    }
    // ---- observables as sparse rows
    c->n_terms = cfg->n_terms;
    c->n_groups = std::max(1, cfg->n_groups);
    std::vector<int32_t> tk, tgroup, rowbase, rowptr;
    std::vector<uint8_t> tsites, mask;
    std::vector<double> val;
    std::vector<int32_t> ratio_term, term_slot;  // terms with an off-diagonal entry
    int64_t so = 0, mo = 0;
    for (int t = 0; t < c->n_terms; t++) {
        const size_t m0 = mask.size();
        int k = cfg->term_k[t];
        if (k < 1 || k > 8) return bad("term with more than 8 sites");
        int g = cfg->term_group ? cfg->term_group[t] : 0;
        if (g < 0 || g >= c->n_groups || (t > 0 && g < tgroup.back())) return bad("term groups must be sorted");
        tk.push_back(k);
        tgroup.push_back(g);
        for (int i = 0; i < 8; i++) {
            int s = i < k ? cfg->term_sites[so + i] : 0;
            if (s < 0 || s >= c->N) return bad("term site out of range");
            tsites.push_back((uint8_t)s);
        }
        so += k;
        int d = 1 << k;
        rowbase.push_back((int32_t)rowptr.size());
        for (int loc = 0; loc < d; loc++) {
            rowptr.push_back((int32_t)mask.size());
            for (int j = 0; j < d; j++) {
                double re = cfg->term_mat[2 * (mo + (int64_t)loc * d + j)];
                double im = cfg->term_mat[2 * (mo + (int64_t)loc * d + j) + 1];
                if (re == 0 && im == 0) continue;
                mask.push_back((uint8_t)(loc ^ j));
                val.push_back(re);
                val.push_back(im);
            }
        }
        rowptr.push_back((int32_t)mask.size());
        mo += (int64_t)d * d;
        const bool off = std::any_of(mask.begin() + m0, mask.end(), [](uint8_t x) { return x != 0; });
        term_slot.push_back(off ? (int32_t)ratio_term.size() : -1);
        if (off) ratio_term.push_back(t);
    }
    c->n_ratio = (int)ratio_term.size();
    // ---- device
    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    int lds = 0;
    HIPCK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    c->lds_limit = (size_t)lds;
    HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        // the static scratch's address on this device, looked up once per
        // device with the device current
        if (dev < 0 || dev >= 64) return bad("device ordinal out of range for the scratch table");
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[dev]) {
            void* sp = nullptr;
            hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !sp)
                return bad(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            g_scratch[dev] = sp;
        }
        c->scratch_dev = g_scratch[dev];
    }
    // timing-only events: no system-scope fence (no cache writeback and
    // invalidation between the kernels they bracket); results are read
    // after a stream synchronisation
    for (auto& e : c->ev) HIPCK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    std::vector<uint8_t> perm8(c->perm.begin(), c->perm.end());
    if (upload(ctx, &c->d_perm, perm8.data(), perm8.size())) return bad(c->err);
    std::vector<int32_t> inv((size_t)c->nsymm * c->N), vbs((size_t)c->nsymm * c->nvb);
    for (int s = 0; s < c->nsymm; s++) {
        for (int f = 0; f < c->N; f++) inv[s * c->N + c->perm[s * c->N + f]] = f;
        for (int i = 0; i < c->nvb; i++) vbs[s * c->nvb + i] = c->perm[s * c->N + c->basis[i]];
    }
    if (upload(ctx, &c->d_invperm, inv.data(), inv.size() * 4)) return bad(c->err);
    if (upload(ctx, &c->d_vbsite, vbs.data(), vbs.size() * 4)) return bad(c->err);
    if (upload(ctx, &c->d_pb, c->pb.data(), c->pb.size() * 4)) return bad(c->err);
    if (c->pf_rows && upload(ctx, &c->d_bs, bs.data(), bs.size() * 2)) return bad(c->err);
    if (upload(ctx, &c->d_tk, tk.data(), tk.size() * 4)) return bad(c->err);
    if (upload(ctx, &c->d_ratio_term, ratio_term.data(), ratio_term.size() * 4)) return bad(c->err);
    if (upload(ctx, &c->d_term_slot, term_slot.data(), term_slot.size() * 4)) return bad(c->err);
    c->tk_max = tk.empty() ? 0 : *std::max_element(tk.begin(), tk.end());
    if (upload(ctx, &c->d_tgroup, tgroup.data(), tgroup.size() * 4)) return bad(c->err);
    if (upload(ctx, &c->d_rowbase, rowbase.data(), rowbase.size() * 4)) return bad(c->err);
    if (upload(ctx, &c->d_rowptr, rowptr.data(), rowptr.size() * 4)) return bad(c->err);
    c->n_rowptr = (int)rowptr.size();
    c->n_ent = (int)mask.size();
    if (upload(ctx, &c->d_tsites, tsites.data(), tsites.size())) return bad(c->err);
    if (upload(ctx, &c->d_mask, mask.data(), mask.size())) return bad(c->err);
    if (upload(ctx, &c->d_val, val.data(), val.size() * 8)) return bad(c->err);
    {
        // k_sample_wave's pair tables: every ordered image (perm_s a, perm_s b)
        // of a proposal bond (a, b) gets an id (at most 255 of them, else
        // bond moves are composed from the single-site tables), and each
        // two-site term that is a proposal bond names it
        const int npb = (int)c->pb.size() / 2;
        std::vector<int16_t> tbond(c->n_terms, -1);
        for (int t = 0; t < c->n_terms; t++) {
            if (tk[t] != 2) continue;
            const int x = tsites[8 * t], y = tsites[8 * t + 1];
            for (int b = 0; b < npb; b++)
                if ((c->pb[2 * b] == x && c->pb[2 * b + 1] == y) || (c->pb[2 * b] == y && c->pb[2 * b + 1] == x)) {
                    tbond[t] = (int16_t)b;
                    break;
                }
        }
        if (tbond.empty()) tbond.push_back(-1);
        if (upload(ctx, &c->d_tbond, tbond.data(), tbond.size() * 2)) return bad(c->err);
        if (c->nsymm > 1 && npb > 0) {
            std::map<std::pair<int, int>, int> ids;
            std::vector<int32_t> pairs;
            std::vector<uint8_t> pmap((size_t)c->nsymm * npb);
            bool fits = true;
            for (int sy = 0; sy < c->nsymm && fits; sy++)
                for (int b = 0; b < npb; b++) {
                    const std::pair<int, int> q{c->perm[sy * c->N + c->pb[2 * b]], c->perm[sy * c->N + c->pb[2 * b + 1]]};
                    auto it = ids.find(q);
                    if (it == ids.end()) {
                        if (ids.size() >= 255) {
                            fits = false;
                            break;
                        }
                        it = ids.emplace(q, (int)ids.size()).first;
                        pairs.push_back(q.first);
                        pairs.push_back(q.second);
                    }
                    pmap[(size_t)sy * npb + b] = (uint8_t)it->second;
                }
            if (fits) {
                c->npair = (int)ids.size();
                if (upload(ctx, &c->d_pmap, pmap.data(), pmap.size())) return bad(c->err);
                if (upload(ctx, &c->d_pairs, pairs.data(), pairs.size() * 4)) return bad(c->err);
            }
        }
    }
    if (dalloc(ctx, &c->d_acc, (size_t)c->chains_cap * 4) || dalloc(ctx, &c->d_prop, (size_t)c->chains_cap * 4))
        return bad(c->err);
    {
        size_t sl = struct_lds_bytes(c->N, c->na);
        c->structured = !c->keep_o && c->nsymm == 1 && c->nvb == c->N && c->pf_rows == 0 &&
                        c->prec == RBMGPU_FP32 && c->nw == 1 && sl > 0 && sl <= c->lds_limit;
        c->ld_s = c->P + c->na + ((c->P + c->na) & 1);
    }
    c->params.assign(c->P, 0.);
    if (build_weight_tables(c)) return bad(c->err);
    if (cfg->learn_hidden > 0) {  // learned proposal
        if (c->prec != RBMGPU_FP32 || c->pf_rows) return bad("the learned proposal needs FP32 and no Pfaffian factor");
        if (cfg->learn_hidden % 16 || cfg->learn_hidden > 512) return bad("learn_hidden: multiple of 16, <= 512");
        if (c->auto_chains) return bad("the learned proposal needs a fixed number of chains");
        if (cfg->cluster_kmax < 1 || cfg->cluster_kmax > 8 || !cfg->clusters) return bad("invalid cluster table");
        c->Hn = cfg->learn_hidden;
        c->Np = (c->N + 15) / 16 * 16;
        c->kmax = cfg->cluster_kmax;
        c->lam = cfg->learn_fraction;
        if (!learn_k(c)) return bad("too many hidden units for the learned sampler");
        if (learn_tables(c, cfg->clusters)) return bad(c->err);
        if (upload(ctx, &c->d_cyc, c->cyc.data(), c->cyc.size())) return bad(c->err);
        if (upload(ctx, &c->d_seedc, c->seedc.data(), c->seedc.size())) return bad(c->err);
        if (upload(ctx, &c->d_seedl, c->seedl.data(), c->seedl.size())) return bad(c->err);
        // initial weights: W1 ~ N(0, 1/N), W2 ~ N(0, 1e-4) (near-uniform seeds), zero biases
        std::mt19937_64 rng(cfg->learn_seed);
        std::normal_distribution<float> d1(0.f, 1.f / std::sqrt((float)c->N)), d2(0.f, 1e-2f);
        std::vector<float> mlp(c->mlp_n(), 0.f);
        for (int64_t i = 0; i < (int64_t)c->Hn * c->N; i++) mlp[i] = d1(rng);
        for (int64_t i = 0; i < (int64_t)c->N * c->Hn; i++) mlp[(int64_t)c->Hn * c->N + c->Hn + i] = d2(rng);
        if (upload(ctx, &c->d_mlp, mlp.data(), mlp.size() * 4)) return bad(c->err);
        std::vector<float> z(2 * mlp.size(), 0.f);
        if (upload(ctx, &c->d_adam, z.data(), z.size() * 4)) return bad(c->err);
        if (dalloc(ctx, &c->d_grad, mlp.size() * 4)) return bad(c->err);
        if (dalloc(ctx, &c->d_W1b, (size_t)c->Hn * c->Np * 2) || dalloc(ctx, &c->d_W2b, (size_t)c->Hn * c->Np * 2) ||
            dalloc(ctx, &c->d_b1p, (size_t)c->Hn * 4) || dalloc(ctx, &c->d_b2p, (size_t)c->Np * 4))
            return bad(c->err);
        if (dalloc(ctx, &c->d_lacc, c->n_chains * 4) || dalloc(ctx, &c->d_lprop, c->n_chains * 4)) return bad(c->err);
        if (learn_pack(c)) return bad(c->err);
        size_t lds = lds_bytes<float>(dims(c, 64 * kMlpChains)) + mlp_lds_bytes(c->Hn, c->Np, c->N, c->kmax);
        if (lds > c->lds_limit) return bad("the learned sampler's tables do not fit LDS");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bad("stream synchronisation failed");
    }
    *out = c;
    return 0;
}

void rbmgpu_destroy(rbmgpu_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->d_W,      c->d_perm,
                    c->d_invperm, c->d_vbsite, c->d_pb,     c->d_tk,    c->d_tgroup, c->d_rowbase, c->d_rowptr,
                    c->d_tsites, c->d_mask,   c->d_val,    c->d_states, c->d_obs,   c->d_T,      c->d_O,
                    c->d_acc,    c->d_prop,   c->d_part,   c->d_varpart, c->d_omean, c->d_ovar,   c->d_eo,     c->d_maxvar,
                    c->d_emom,   c->d_epart,  c->d_x,      c->d_b,      c->d_r,     c->d_p,      c->d_Ap,     c->d_st,
                    c->d_sxpart, c->d_sxpart2, c->d_evst,   c->d_evout,  c->d_evobs, c->d_bs,
                    c->d_Opf,    c->d_pdot,   c->d_cyc,    c->d_seedc, c->d_seedl,  c->d_W1b,    c->d_W2b,
                    c->d_b1p,    c->d_b2p,    c->d_mlp,    c->d_adam,  c->d_grad,   c->d_rec_st, c->d_rec_seed,
                    c->d_rec_k,  c->d_rec_a,  c->d_lacc,   c->d_lprop, c->d_tr,     c->d_tst,    c->d_tseed,
                    c->d_tw,     c->d_tpart,     c->d_ratio_term, c->d_term_slot, c->d_params, c->d_par, c->d_stats, c->d_Smat, c->d_gpart, c->d_w, c->d_wsx, c->d_lp,
                    c->d_wred, c->d_pmap, c->d_pairs, c->d_tbond};
    drop_graphs(c);
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_par) (void)hipHostFree(c->h_par);
    if (c->h_stats) (void)hipHostFree(c->h_stats);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int rbmgpu_n_params(const rbmgpu_ctx* c, int64_t* n, int64_t* nn) {
    if (!c) return -1;
    if (n) *n = c->P;
    if (nn) *nn = c->P_rbm;
    return 0;
}

int rbmgpu_n_chains(const rbmgpu_ctx* c, int64_t* n) {
    if (!c || !n) return -1;
    *n = c->n_chains;
    return 0;
}

int rbmgpu_launch_shape(const rbmgpu_ctx* c, int32_t* o) {
    if (!c) return -1;
    o[0] = c->shape.L;
    o[1] = c->shape.K;
    o[2] = c->block;
    o[3] = (c->structured ? 1 : 0) | (c->shape.wave ? 2 : 0);
    return 0;
}

}  // extern "C"

namespace {
// the FP64 parameters on the device follow the host copy
int push_params(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (int rc = build_weight_tables(c)) return rc;
    if (!c->d_params)
        if (int rc = dalloc(ctx, &c->d_params, c->P * 2 * sizeof(double))) return rc;
    HIPCK(hipStreamSynchronize(c->stream));
    HIPCK(hipMemcpy(c->d_params, c->params.data(), c->P * 16, hipMemcpyHostToDevice));
    c->params_dev = false;
    return 0;
}
// the host copy from the device after device-side updates
int pull_params(rbmgpu_ctx* c) {
    rbmgpu_ctx* ctx = c;
    if (!c->params_dev) return 0;
    HIPCK(hipStreamSynchronize(c->stream));
    HIPCK(hipMemcpy(c->params.data(), c->d_params, c->P * 16, hipMemcpyDeviceToHost));
    c->params_dev = false;
    return 0;
}
}  // namespace

extern "C" {

int rbmgpu_set_params(rbmgpu_ctx* c, const double* p) {
    if (!c || !p) return -1;
    for (int64_t i = 0; i < c->P; i++) c->params[i] = zd(p[2 * i], p[2 * i + 1]);
    return push_params(c);
}

int rbmgpu_get_params(const rbmgpu_ctx* c, double* p) {
    if (!c || !p) return -1;
    if (c->params_dev) {
        if (hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(p, c->d_params, c->P * 16, hipMemcpyDeviceToHost) != hipSuccess)
            return -2;
        return 0;
    }
    for (int64_t i = 0; i < c->P; i++) {
        p[2 * i] = c->params[i].real();
        p[2 * i + 1] = c->params[i].imag();
    }
    return 0;
}

int rbmgpu_update_weights(rbmgpu_ctx* c, const double* dw) {
    if (!c || !dw) return -1;
    if (int rc = pull_params(c)) return rc;
    for (int64_t i = 0; i < c->P; i++) c->params[i] -= zd(dw[2 * i], dw[2 * i + 1]);
    return push_params(c);
}

int rbmgpu_train_epoch(rbmgpu_ctx* c, uint32_t epoch, const rbmgpu_sr_params* sp, rbmgpu_epoch_stats* out) {
    if (!c || !sp) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->d_params) return fail(c, "rbmgpu_train_epoch: no parameters (rbmgpu_set_params first)");
    if (int rc = autotune(c)) return rc;
    if (int rc = ensure_samples(c, c->n_samples)) return rc;
    if (int rc = ensure_cg(c)) return rc;
    if (int rc = ensure_train(c)) return rc;
    // cg_solver.cpp:48-60: max_iter = min(2P, max_iter), 0 -> 2P; rtol 0 -> 1e-16
    int64_t mi = std::min<int64_t>(2 * c->P, sp->max_iter);
    if (mi <= 0) mi = 2 * c->P;
    if (c->train_dense < 0) c->train_dense = decide_train_dense(c, mi);
    if (c->structured) c->train_dense = 0;  // no O matrix to build S from
    DevPar* hp = (DevPar*)c->h_par;
    hp->epoch = epoch;
    hp->pad = 0;
    hp->max_iter = mi;
    hp->reg[0] = sp->reg1;
    hp->reg[1] = sp->reg2;
    hp->reg[2] = sp->reg1_delta;
    hp->lr = sp->lr;
    hp->rtol = sp->rtol == 0. ? 1e-16 : sp->rtol;
    HIPCK(hipMemcpyAsync(c->d_par, c->h_par, sizeof(DevPar), hipMemcpyHostToDevice, c->stream));
    // CG launches captured: the last solve's count rounded up to 8 (launches
    // past convergence return at once); an unfinished solve continues below
    const int64_t budget = std::min<int64_t>(mi, std::max<int64_t>(8, (c->train_itn + 2 + 7) / 8 * 8));
    auto t0 = std::chrono::steady_clock::now();
    if (sp->graph) {
        if (c->graph_gen != c->gen) drop_graphs(c);
        hipGraphExec_t ex = nullptr;
        for (auto& g : c->graphs)
            if (g.first == budget) ex = g.second;
        if (ex) {
            HIPCK(hipGraphLaunch(ex, c->stream));
        } else {
            // this epoch runs eagerly (allocating what it needs), then the
            // same body is captured for the following epochs
            if (int rc = train_body(c, budget)) return rc;
            HIPCK(hipStreamSynchronize(c->stream));
            if (c->graph_gen != c->gen) drop_graphs(c);
            const uint64_t gen0 = c->gen;
            hipGraph_t g = nullptr;
            c->capturing = true;
            hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
            int rc = e == hipSuccess ? train_body(c, budget) : 0;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            c->capturing = false;
            if (e != hipSuccess) return fail(c, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e2 != hipSuccess) return fail(c, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
            if (c->gen != gen0) {
                (void)hipGraphDestroy(g);
                return fail(c, "device buffers changed during graph capture");
            }
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(c, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            c->graphs.push_back({budget, ex});
            c->graph_gen = c->gen;
        }
    } else {
        if (int rc = train_body(c, budget)) return rc;
    }
    DevStats* hs = (DevStats*)c->h_stats;
    HIPCK(hipMemcpyAsync(c->h_stats, c->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    while (hs->pending) {
        if (int rc = train_more(c, 8)) return rc;
        HIPCK(hipMemcpyAsync(c->h_stats, c->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->train_itn = hs->itn;
    c->cg_last_itn = hs->itn;
    c->moments_valid = true;
    if (hs->ok) c->params_dev = true;
    if (out) {
        out->e_mean[0] = hs->e[0];
        out->e_mean[1] = hs->e[1];
        out->e_var = hs->e[2];
        out->e_var_binned = hs->e[3];
        out->e_tau = hs->e[4];
        out->e_std = hs->e[5];
        int64_t nb = 50;
        while (c->S % nb != 0) nb--;
        out->n_bins = nb;
        out->acceptance = hs->acc;
        out->cg_itn = hs->itn;
        out->cg_rs = hs->rs;
        out->max_dw = hs->max_dw;
        out->updated = hs->ok;
        out->graph_replayed = sp->graph && !c->graphs.empty();
        out->ms = ms;
        out->explicit_s = c->train_dense == 1 ? 1 : 0;
    }
    return 0;
}

int rbmgpu_sample_epoch(rbmgpu_ctx* c, uint32_t epoch) {
    if (!c) return -1;
    if (int rc = autotune(c)) return rc;
    if (int rc = ensure_samples(c, c->n_samples)) return rc;
    return sample_body(c, epoch);
}

int rbmgpu_full_epoch(rbmgpu_ctx* c, int32_t parallel_bits) {
    if (!c) return -1;
    rbmgpu_ctx* ctx = c;
    if (c->N > 30) return fail(c, "exact enumeration: at most 30 spins here (2^N samples on the device)");
    if (c->structured) return fail(c, "exact enumeration needs the O matrix (keep_o)");
    const int bp = std::max(0, std::min((int)parallel_bits, c->N));
    const int64_t S = (int64_t)1 << c->N;
    if (int rc = ensure_samples(c, S)) return rc;
    const int nb = (int)std::min<int64_t>((S + 255) / 256, 1024);
    if (S > c->w_cap) {
        if (int rc = dalloc(ctx, &c->d_w, S * sizeof(double))) return rc;
        if (int rc = dalloc(ctx, &c->d_wsx, S * sizeof(double))) return rc;
        if (int rc = dalloc(ctx, &c->d_lp, S * 2 * sizeof(double))) return rc;
        if (!c->d_wred)
            if (int rc = dalloc(ctx, &c->d_wred, 2 * 1024 * sizeof(double))) return rc;
        c->w_cap = S;
    }
    hipLaunchKernelGGL(k_gray_states, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, c->stream, c->d_states, S,
                       c->nw, bp, c->N);
    HIPCK(hipGetLastError());
    if (int rc = zero_async(c, c->d_obs, S * 2 * c->n_groups * sizeof(double))) return rc;
    if (int rc = record(c, 0)) return rc;
    if (int rc = eval(c, c->d_states, S, c->d_lp, c->d_obs, c->d_T, nullptr, nullptr, 0)) return rc;
    if (int rc = record(c, 1)) return rc;
    hipLaunchKernelGGL(k_lp_max, dim3(nb), dim3(256), 0, c->stream, c->d_lp, S, c->d_wred);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_weights, dim3(nb), dim3(256), 0, c->stream, c->d_lp, S, c->d_wred, nb, c->d_w,
                       c->d_wred + 1024);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_wsx, dim3(nb), dim3(256), 0, c->stream, c->d_w, S, c->d_wred + 1024, nb, c->d_wsx);
    HIPCK(hipGetLastError());
    c->S = S;
    c->have_chains = false;
    c->weighted = true;
    int rc = c->prec == RBMGPU_FP32 ? run_build_O<float>(c) : run_build_O<double>(c);
    if (rc) return rc;
    if (int rc2 = record(c, 2)) return rc2;
    HIPCK(hipStreamSynchronize(c->stream));
    c->moments_valid = false;
    return 0;
}

int rbmgpu_full_weights(const rbmgpu_ctx* c, double* logpsi, double* w) {
    if (!c || !c->weighted) return -1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -2;
    if (logpsi && hipMemcpy(logpsi, c->d_lp, c->S * 16, hipMemcpyDeviceToHost) != hipSuccess) return -2;
    if (w && hipMemcpy(w, c->d_w, c->S * 8, hipMemcpyDeviceToHost) != hipSuccess) return -2;
    return 0;
}

int rbmgpu_load_samples(rbmgpu_ctx* c, const uint64_t* states, int64_t n) {
    if (!c || !states || n < 1) return -1;
    rbmgpu_ctx* ctx = c;
    if (int rc = ensure_samples(c, n)) return rc;
    c->weighted = false;
    HIPCK(hipMemcpyAsync(c->d_states, states, n * c->nw * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
    if (int rc = zero_async(c, c->d_obs, n * 2 * c->n_groups * sizeof(double))) return rc;
    HIPCK(hipEventRecord(c->ev[0], c->stream));
    if (int rc = eval(c, c->d_states, n, nullptr, c->d_obs, c->d_T, nullptr, nullptr, 0)) return rc;
    HIPCK(hipEventRecord(c->ev[1], c->stream));
    c->S = n;
    c->have_chains = false;
    int rc = c->prec == RBMGPU_FP32 ? run_build_O<float>(c) : run_build_O<double>(c);
    if (rc) return rc;
    HIPCK(hipEventRecord(c->ev[2], c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    c->moments_valid = false;
    return 0;
}

int rbmgpu_reduce(rbmgpu_ctx* c, rbmgpu_moments* m, double* o_mean, double* o_var, double* eo,
                  double* gmeans) {
    if (!c) return -1;
    rbmgpu_ctx* ctx = c;
    if (int rc = compute_moments(c)) return rc;
    std::vector<double> e(16 + 2 * c->n_groups);
    HIPCK(hipMemcpyAsync(e.data(), c->d_emom, e.size() * 8, hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> acc, prop;
    if (c->have_chains) {
        acc.resize(c->n_chains);
        prop.resize(c->n_chains);
        HIPCK(hipMemcpyAsync(acc.data(), c->d_acc, c->n_chains * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipMemcpyAsync(prop.data(), c->d_prop, c->n_chains * 4, hipMemcpyDeviceToHost, c->stream));
    }
    if (o_mean) HIPCK(hipMemcpyAsync(o_mean, c->d_omean, c->P * 16, hipMemcpyDeviceToHost, c->stream));
    if (o_var) HIPCK(hipMemcpyAsync(o_var, c->d_ovar, c->P * 8, hipMemcpyDeviceToHost, c->stream));
    if (eo) HIPCK(hipMemcpyAsync(eo, c->d_eo, c->P * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (m) {
        m->e_mean[0] = e[0];
        m->e_mean[1] = e[1];
        m->e_var = e[2];
        m->e_var_binned = e[3];
        m->e_tau = e[4];
        m->e_std = e[5];
        int64_t nb = 50;
        while (c->S % nb != 0) nb--;
        m->n_bins = nb;
        double a = 0;
        for (int i = 0; i < (int)acc.size(); i++) a += prop[i] ? (double)acc[i] / prop[i] : 0.0;
        m->acceptance = acc.empty() ? 0.0 : a / c->n_chains;
    }
    if (gmeans) std::memcpy(gmeans, e.data() + 16, 2 * c->n_groups * sizeof(double));
    c->times[0] = elapsed(c, 0, 1);
    c->times[1] = elapsed(c, 1, 2);
    c->times[2] = elapsed(c, 4, 5);
    return 0;
}

int rbmgpu_apply_S(rbmgpu_ctx* c, const double* x, double* y, double r1, double r2, double rd) {
    if (!c || !x || !y) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->moments_valid)
        if (int rc = compute_moments(c)) return rc;
    if (int rc = ensure_cg(c)) return rc;
    if (int rc = ensure_sx(c)) return rc;
    HIPCK(hipMemcpyAsync(c->d_p, x, c->P * 16, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipEventRecord(c->ev[6], c->stream));
    if (int rc = sx(c, c->d_p, c->d_Ap, r1, r2, rd, nullptr)) return rc;
    HIPCK(hipEventRecord(c->ev[7], c->stream));
    HIPCK(hipMemcpyAsync(y, c->d_Ap, c->P * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    c->times[3] = elapsed(c, 6, 7);
    return 0;
}

int rbmgpu_apply_S_noalloc(rbmgpu_ctx* c, const double* d_x, double* d_y, double r1, double r2, double rd) {
    if (!c || !d_x || !d_y) return -1;
    if (!c->scratch_dev) return fail(c, "rbmgpu_apply_S_noalloc: no scratch address for this device");
    if (!c->moments_valid)
        return fail(c, "rbmgpu_apply_S_noalloc: no moments of the current sample set (call rbmgpu_reduce first)");
    int64_t blocks, bytes;
    if (int rc = sx_layout(c, &blocks, &bytes)) return rc;
    if (bytes > kScratchBytes)
        return fail(c, "rbmgpu_apply_S_noalloc: the S.x block partials need " + std::to_string(bytes) +
                           " bytes, more than the static scratch's " + std::to_string(kScratchBytes));
    c->sx_blocks = blocks;
    return sx(c, d_x, d_y, r1, r2, rd, nullptr, nullptr, false, c->scratch_dev);
}

int rbmgpu_learn_set_fraction(rbmgpu_ctx* c, double fraction) {
    if (!c || fraction < 0 || fraction > 1) return -1;
    c->gen++;  // the sampler kernel and its arguments change
    if (!c->Hn && fraction > 0) return fail(c, "no proposal network (learn_hidden = 0)");
    c->lam = fraction;
    return 0;
}

int rbmgpu_learn_weights(rbmgpu_ctx* c, int32_t set, double* params) {
    if (!c || !params) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->Hn) return fail(c, "no proposal network (learn_hidden = 0)");
    std::vector<float> f(c->mlp_n());
    if (set) {
        for (size_t i = 0; i < f.size(); i++) f[i] = (float)params[i];
        HIPCK(hipMemcpyAsync(c->d_mlp, f.data(), f.size() * 4, hipMemcpyHostToDevice, c->stream));
        if (int rc = learn_pack(c)) return rc;
        HIPCK(hipStreamSynchronize(c->stream));
    } else {
        HIPCK(hipMemcpyAsync(f.data(), c->d_mlp, f.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < f.size(); i++) params[i] = f[i];
    }
    return 0;
}

int rbmgpu_learn_logits(rbmgpu_ctx* c, const uint64_t* states, int64_t n, double* z) {
    if (!c || !states || !z || n < 1) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->Hn) return fail(c, "no proposal network (learn_hidden = 0)");
    if (int rc = ensure_eval(c, n)) return rc;
    float* dz = nullptr;
    HIPCK(hipMalloc((void**)&dz, (size_t)n * c->N * 4));
    HIPCK(hipMemcpyAsync(c->d_evst, states, n * c->nw * 8, hipMemcpyHostToDevice, c->stream));
    LearnArgs a = learn_args(c, 0);
    hipError_t e = launch_mlp_logits(a, c->d_evst, n, dz, c->stream);
    std::vector<float> h((size_t)n * c->N);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dz, h.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dz);
    HIPCK(e);
    for (size_t i = 0; i < h.size(); i++) z[i] = h[i];
    return 0;
}

int rbmgpu_learn_grad(rbmgpu_ctx* c, const uint64_t* states, const int32_t* seeds, const double* w, int64_t n,
                      double beta, double* grads, double* loss) {
    if (!c || !states || !seeds || !w || !grads || n < 1) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->Hn) return fail(c, "no proposal network (learn_hidden = 0)");
    if (int rc = ensure_train(c, n)) return rc;
    std::vector<int16_t> sd(n);
    std::vector<float> wf(n);
    for (int64_t i = 0; i < n; i++) {
        if (seeds[i] < -1 || seeds[i] >= c->N) return fail(c, "seed out of range");
        sd[i] = (int16_t)seeds[i];
        wf[i] = (float)w[i];
    }
    HIPCK(hipMemcpyAsync(c->d_tst, states, n * c->nw * 8, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->d_tseed, sd.data(), n * 2, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->d_tw, wf.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    if (int rc = learn_grad(c, n, beta, loss)) return rc;
    std::vector<float> g(c->mlp_n());
    HIPCK(hipMemcpyAsync(g.data(), c->d_grad, g.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < g.size(); i++) grads[i] = g[i];
    return 0;
}

int rbmgpu_learn_train(rbmgpu_ctx* c, double lr, double beta, int64_t batch, double* out4) {
    if (!c || batch < 1) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->Hn) return fail(c, "no proposal network (learn_hidden = 0)");
    if (!c->learned_epoch || c->S == 0) return fail(c, "no records: sample an epoch with the learned sampler first");
    const int64_t S = c->S;
    std::vector<uint64_t> st(S * c->nw);
    std::vector<int16_t> sd(S);
    std::vector<uint8_t> kk(S);
    std::vector<float> aa(S);
    std::vector<uint32_t> lacc(c->n_chains), lprop(c->n_chains);
    HIPCK(hipMemcpyAsync(st.data(), c->d_rec_st, st.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(sd.data(), c->d_rec_seed, S * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(kk.data(), c->d_rec_k, S, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(aa.data(), c->d_rec_a, S * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(lacc.data(), c->d_lacc, c->n_chains * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(lprop.data(), c->d_lprop, c->n_chains * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    // batch: valid records spread evenly over the sample set
    std::vector<int64_t> idx;
    for (int64_t r = 0; r < S; r++)
        if (sd[r] >= 0) idx.push_back(r);
    double np = 0, na = 0;
    for (int i = 0; i < c->n_chains; i++) np += lprop[i], na += lacc[i];
    if (out4) {
        out4[0] = 0;
        out4[1] = 0;
        out4[2] = np;
        out4[3] = na;
    }
    if (idx.empty()) return 0;
    int64_t B = std::min<int64_t>(batch, (int64_t)idx.size());
    std::vector<uint64_t> bst(B * c->nw);
    std::vector<int16_t> bsd(B);
    std::vector<float> bw(B);
    double mean = 0;
    for (int64_t b = 0; b < B; b++) {
        int64_t r = idx[(size_t)((double)b * idx.size() / B)];
        for (int q = 0; q < c->nw; q++) bst[b * c->nw + q] = st[r * c->nw + q];
        bsd[b] = sd[r];
        bw[b] = aa[r] * kk[r];  // reward: expected number of spins moved
        mean += bw[b];
    }
    mean /= B;
    for (auto& x : bw) x = (float)(x - mean);
    if (int rc = ensure_train(c, B)) return rc;
    HIPCK(hipMemcpyAsync(c->d_tst, bst.data(), bst.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->d_tseed, bsd.data(), B * 2, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->d_tw, bw.data(), B * 4, hipMemcpyHostToDevice, c->stream));
    double loss = 0;
    if (int rc = learn_grad(c, B, beta, &loss)) return rc;
    c->learn_t++;
    const int64_t n = c->mlp_n();
    float c1 = 1.f - std::pow(0.9f, (float)c->learn_t), c2 = 1.f - std::pow(0.999f, (float)c->learn_t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_mlp, c->d_grad,
                       c->d_adam, c->d_adam + n, n, (float)lr, c1, c2);
    HIPCK(hipGetLastError());
    if (int rc = learn_pack(c)) return rc;
    HIPCK(hipStreamSynchronize(c->stream));
    if (out4) {
        out4[0] = loss;
        out4[1] = mean;
    }
    return 0;
}

int rbmgpu_time_sx(rbmgpu_ctx* c, int32_t reps, double* ms) {
    if (!c || !ms || reps < 1) return -1;
    rbmgpu_ctx* ctx = c;
    if (!c->moments_valid)
        if (int rc = compute_moments(c)) return rc;
    if (int rc = ensure_cg(c)) return rc;
    if (int rc = ensure_sx(c)) return rc;
    // the last CG direction (or zeros) as the operand; only the main kernel runs
    HIPCK(hipEventRecord(c->ev[8], c->stream));
    for (int i = 0; i < reps; i++) {
        if (c->structured) {
            if (int rc = run_struct(c, 0, c->d_p, nullptr)) return rc;
        } else if (c->prec == RBMGPU_FP32) {
            SxArgs<float> a = sx_args<float>(c, c->d_p, c->d_Ap, 0, 0, 0, nullptr);
            HIPCK(launch_sx<float>(a, (int)c->sx_blocks, c->stream));
        } else {
            SxArgs<double> a = sx_args<double>(c, c->d_p, c->d_Ap, 0, 0, 0, nullptr);
            HIPCK(launch_sx<double>(a, (int)c->sx_blocks, c->stream));
        }
    }
    HIPCK(hipEventRecord(c->ev[9], c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    *ms = elapsed(c, 8, 9) / reps;
    c->times[5] = *ms;
    return 0;
}

int rbmgpu_sr_solve(rbmgpu_ctx* c, const double* b, double* x, double r1, double r2, double rd,
                    int64_t max_iter, double rtol, int64_t* itn, double* rs) {
    return rbmgpu_sr_solve_ex(c, RBMGPU_SOLVE_CG, b, x, r1, r2, rd, max_iter, rtol, itn, rs);
}

int rbmgpu_build_S(rbmgpu_ctx* c, double r1, double r2, double rd, double* S) {
    if (!c || !S) return -1;
    rbmgpu_ctx* ctx = c;
    if (int rc = build_S(c, r1, r2, rd, 0)) return rc;
    HIPCK(hipMemcpyAsync(S, c->d_Smat, c->P * c->P * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return 0;
}

int rbmgpu_sr_solve_ex(rbmgpu_ctx* c, int32_t method, const double* b, double* x, double r1, double r2, double rd,
                       int64_t max_iter, double rtol, int64_t* itn, double* rs) {
    if (!c || !b || !x) return -1;
    if (method < RBMGPU_SOLVE_CG || method > RBMGPU_SOLVE_DIRECT) return fail(c, "unknown SR solve method");
    rbmgpu_ctx* ctx = c;
    if (!c->moments_valid)
        if (int rc = compute_moments(c)) return rc;
    if (int rc = ensure_cg(c)) return rc;
    if (int rc = ensure_sx(c)) return rc;
    if (method == RBMGPU_SOLVE_DIRECT) {
        if (int rc = build_S(c, r1, r2, rd, 1)) return rc;
        std::vector<zd> Sm((size_t)c->P * c->P), bb(c->P), xx(c->P);
        HIPCK(hipMemcpyAsync(Sm.data(), c->d_Smat, c->P * c->P * 16, hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < c->P; i++) bb[i] = zd(b[2 * i], b[2 * i + 1]);
        if (int rc = dense_solve(c, Sm, bb, xx)) return rc;
        double r2s = 0;  // |S x - b|^2 of the solution
        for (int64_t i = 0; i < c->P; i++) {
            zd acc = -bb[i];
            for (int64_t j = 0; j < c->P; j++) acc += Sm[i + j * c->P] * xx[j];
            r2s += std::norm(acc);
        }
        for (int64_t i = 0; i < c->P; i++) {
            x[2 * i] = xx[i].real();
            x[2 * i + 1] = xx[i].imag();
        }
        if (itn) *itn = 0;
        if (rs) *rs = r2s;
        return 0;
    }
    const bool dense = method == RBMGPU_SOLVE_CG_SINGLE;
    if (dense)
        if (int rc = build_S(c, r1, r2, rd, 0)) return rc;
    // cg_solver.cpp:48-60: max_iter = min(2P, max_iter), 0 -> 2P; rtol 0 -> 1e-16
    int64_t mi = std::min<int64_t>(2 * c->P, max_iter);
    if (mi == 0) mi = 2 * c->P;
    if (rtol == 0.) rtol = 1e-16;
    // y = A x (the regularised S), CG step appended when cg
    auto matvec = [&](const double* xin, double* yout, const CgStatus* st, bool cg) -> int {
        if (!dense) return sx(c, xin, yout, r1, r2, rd, st, cg ? c->d_pdot : nullptr, cg);
        if (cg) return dense_cg_step(c);  // xin = p, yout = Ap
        HIPCK(launch_dense_mv(c->d_Smat, xin, yout, (int)c->P, nullptr, st, c->stream));
        return 0;
    };
    HIPCK(hipMemcpyAsync(c->d_b, b, c->P * 16, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->d_x, x, c->P * 16, hipMemcpyHostToDevice, c->stream));
    if (int rc = matvec(c->d_x, c->d_Ap, nullptr, false)) return rc;
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1024), 0, c->stream, c->d_b, c->d_Ap, c->d_r, c->d_p, (int)c->P,
                       c->d_st, mi, rtol, (const DevPar*)nullptr);
    HIPCK(hipGetLastError());
    CgStatus st{};
    int64_t launched = 0;
    // the loop is queued in batches with one status read-back each: the
    // first as long as the last solve (warm-started solves of consecutive
    // epochs converge in similar counts), then kCgBatch more at a time;
    // launches queued past convergence return at once
    constexpr int64_t kCgBatch = 8;
    int64_t nb = std::max<int64_t>(kCgBatch, c->cg_last_itn + 1);
    HIPCK(hipEventRecord(c->ev[12], c->stream));
    while (launched < mi) {
        nb = std::min<int64_t>(nb, mi - launched);
        for (int64_t i = 0; i < nb; i++)
            if (int rc = matvec(c->d_p, c->d_Ap, c->d_st, true)) return rc;
        HIPCK(hipMemcpyAsync(&st, c->d_st, sizeof(CgStatus), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        launched += nb;
        nb = kCgBatch;
        if (st.done) break;
    }
    HIPCK(hipEventRecord(c->ev[13], c->stream));
    HIPCK(hipMemcpyAsync(x, c->d_x, c->P * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (itn) *itn = st.itn;
    if (rs) *rs = st.rs;
    c->cg_last_itn = st.itn;
    // CG iterations that did work: itn (converged) or max_iter
    if (st.itn > 0) c->times[3] = elapsed(c, 12, 13) / (double)st.itn;
    return 0;
}

int rbmgpu_eval_logpsi(rbmgpu_ctx* c, const uint64_t* states, int64_t n, double* out) {
    if (!c || !states || !out || n < 1) return -1;
    rbmgpu_ctx* ctx = c;
    if (int rc = ensure_eval(c, n)) return rc;
    HIPCK(hipMemcpyAsync(c->d_evst, states, n * c->nw * 8, hipMemcpyHostToDevice, c->stream));
    if (int rc = eval(c, c->d_evst, n, c->d_evout, nullptr, nullptr, nullptr, nullptr, 0)) return rc;
    HIPCK(hipMemcpyAsync(out, c->d_evout, n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return 0;
}

int rbmgpu_psi_ratio(rbmgpu_ctx* c, const uint64_t* state, const int32_t* flips, int32_t nf, double* ratio) {
    if (!c || !state || !ratio || nf < 0 || nf > 8) return -1;
    rbmgpu_ctx* ctx = c;
    for (int i = 0; i < nf; i++)
        if (flips[i] < 0 || flips[i] >= c->N) return fail(c, "flip site out of range");
    if (nf == 0) {
        ratio[0] = 1;
        ratio[1] = 0;
        return 0;
    }
    if (int rc = ensure_eval(c, 1)) return rc;
    HIPCK(hipMemcpyAsync(c->d_evst, state, c->nw * 8, hipMemcpyHostToDevice, c->stream));
    if (int rc = eval(c, c->d_evst, 1, nullptr, nullptr, nullptr, c->d_evout, flips, nf)) return rc;
    HIPCK(hipMemcpyAsync(ratio, c->d_evout, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return 0;
}

int rbmgpu_get_samples(const rbmgpu_ctx* cc, uint64_t* states, double* obs, double* O, double* chain_acc) {
    rbmgpu_ctx* c = const_cast<rbmgpu_ctx*>(cc);
    if (!c) return -1;
    rbmgpu_ctx* ctx = c;
    HIPCK(hipStreamSynchronize(c->stream));
    if (states) HIPCK(hipMemcpy(states, c->d_states, c->S * c->nw * 8, hipMemcpyDeviceToHost));
    if (obs) HIPCK(hipMemcpy(obs, c->d_obs, c->S * 2 * c->n_groups * 8, hipMemcpyDeviceToHost));
    if (O) {
        if (c->structured) return fail(c, "the O matrix is not materialised in structured mode (keep_o = 0)");
        if (c->prec == RBMGPU_FP64) {
            HIPCK(hipMemcpy2D(O, c->P * 16, c->d_O, c->ld * 16, c->P * 16, c->S, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(c->S * c->ld * 2);
            HIPCK(hipMemcpy(f.data(), c->d_O, f.size() * 4, hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < c->S; k++)
                for (int64_t i = 0; i < 2 * c->P; i++) O[k * 2 * c->P + i] = f[k * 2 * c->ld + i];
        }
    }
    if (chain_acc) {
        if (!c->have_chains) return fail(c, "no sampled chains");
        std::vector<uint32_t> a(c->n_chains), p(c->n_chains);
        HIPCK(hipMemcpy(a.data(), c->d_acc, c->n_chains * 4, hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(p.data(), c->d_prop, c->n_chains * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < c->n_chains; i++) chain_acc[i] = p[i] ? (double)a[i] / p[i] : 0.0;
    }
    return 0;
}

int rbmgpu_n_samples(const rbmgpu_ctx* c, int64_t* n) {
    if (!c || !n) return -1;
    *n = c->S;
    return 0;
}

int rbmgpu_kernel_times(const rbmgpu_ctx* c, double* ms6) {
    if (!c || !ms6) return -1;
    for (int i = 0; i < 6; i++) ms6[i] = c->times[i];
    return 0;
}

int rbmgpu_synchronize(rbmgpu_ctx* c) {
    if (!c) return -1;
    rbmgpu_ctx* ctx = c;
    HIPCK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
