// kern_impl.hpp -- gfx950 kernels of the VMC epoch, templated on the real
// type R (float / double).  Included by kern_f32.hip and kern_f64.hip, which
// instantiate the launchers for one precision each.
//
// Chain layout.  A Markov chain is owned by an aligned group of L lanes of
// a 64-wide wavefront; hidden unit h = lane + u*L (u < K) lives in that
// lane's registers as theta_h and tanh(theta_h).  The spin state (<= 128
// bits) is replicated in every lane of the group.  All lanes of a group draw
// the same Philox numbers and reduce with xor butterflies, so the group's
// control flow is uniform without any broadcast.
//
// Ratio identity.  Flipping the spins F changes theta_h by
// delta_h = sum_{f in F} -2 sigma_f W_{k(f,h)}, and
//   cosh(theta + delta) / cosh(theta) = cosh(delta) + tanh(theta) sinh(delta).
// cosh/sinh of a single flip's delta are the weight-only tables
// C = cosh(2W), S = sinh(2W) (with sinh odd in sigma); several flips are
// composed with the addition theorems.  A proposal therefore costs a few
// complex multiply-adds per hidden unit and no transcendental; tanh(theta)
// is recomputed only for accepted moves.
#pragma once
#include "dev_common.hpp"
#include "kern_api.hpp"

namespace rbmgpu {

__device__ inline int bitof(const uint64_t* st, int i) { return (int)((st[i >> 6] >> (i & 63)) & 1); }

// Pointers into LDS are derived from the dynamic shared-memory base only
// (never selected against a global pointer), so the compiler emits ds_read
// for them; the tables that may live in either place have one pointer of
// each kind and the code branches on the (uniform) staging flag.
template <typename R>
struct Lds {
    const cplx<R>* C;
    const cplx<R>* S;
    const cplx<R>* Wl;  // W in LDS (d.w_lds)
    const cplx<R>* Wg;  // W in memory
    const uint8_t* P;   // symmetry permutations (LDS; n_symm > 1)
    const cplx<R>* rb;  // ratio bias per site
    const int32_t* pb;  // proposal bonds (a, b)
    Terms Tl;           // local operators: LDS copies (d.t_lds)
    Terms Tg;           //                  in memory
    PfLds<R> pf;        // Pfaffian tables and this wave's scratch (d.pf_rows > 0)
};

template <typename T>
__device__ inline T* lds_copy(char*& at, const T* src, size_t n) {
    T* dst = reinterpret_cast<T*>(at);
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
    at += n * sizeof(T);
    return dst;
}

// stage the weight tables, visible biases, proposal bonds, local-operator
// tables and symmetry permutations into LDS (layout of lds_bytes)
template <typename R>
__device__ inline Lds<R> stage(const Dims& d, const Tabs<R>& t, const Terms& T, const int32_t* pb,
                               char* smem) {
    const size_t nw = (size_t)d.N * d.n_alpha;
    char* at = smem;
    Lds<R> l;
    l.C = lds_copy(at, t.C, nw);
    l.S = lds_copy(at, t.S, nw);
    l.Wg = t.W;
    l.Wl = d.w_lds ? lds_copy(at, t.W, nw) : reinterpret_cast<const cplx<R>*>(at);
    l.rb = lds_copy(at, t.rb, d.N);
    l.pf = {};
    if (d.pf_rows) {
        l.pf.fs = lds_copy(at, t.fs, 4 * (size_t)d.pf_rows);
        const int sw = pf_scratch<R>(d.N, d.pf_rows);
        cplx<R>* scr = reinterpret_cast<cplx<R>*>(at) + (size_t)(threadIdx.x >> 6) * sw;
        at += (size_t)d.waves * sw * sizeof(cplx<R>);
        l.pf.N = d.N;
        l.pf.rows = d.pf_rows;
        l.pf.X = scr;
        l.pf.Xd = reinterpret_cast<cplx<double>*>(scr);
        l.pf.D = scr + pf_xsize<R>(d.N, d.pf_rows);
        l.pf.G = l.pf.D + 8 * 32;
        l.pf.F = reinterpret_cast<int*>(l.pf.G + 16 * 16);
    }
    l.Tg = T;
    l.Tl = T;
    if (d.t_lds) {
        l.Tl.val = lds_copy(at, T.val, 2 * (size_t)d.n_ent);
        l.Tl.rowptr = lds_copy(at, T.rowptr, d.n_rowptr);
        l.Tl.k = lds_copy(at, T.k, d.n_terms);
        l.Tl.group = lds_copy(at, T.group, d.n_terms);
        l.Tl.rowbase = lds_copy(at, T.rowbase, d.n_terms);
    }
    l.pb = lds_copy(at, pb, 2 * (size_t)d.npb);
    if (d.pf_rows) l.pf.bs = lds_copy(at, t.bs, (size_t)d.N * d.N);
    if (d.t_lds) {
        l.Tl.sites = lds_copy(at, T.sites, 8 * (size_t)d.n_terms);
        l.Tl.mask = lds_copy(at, T.mask, d.n_ent);
    }
    l.P = d.n_symm > 1 ? lds_copy(at, t.perm, (size_t)d.n_symm * d.N) : reinterpret_cast<const uint8_t*>(at);
    __syncthreads();
    return l;
}

__device__ inline int kidx(const Dims& d, const uint8_t* P, int s, int f) {
    return d.n_symm == 1 ? f : (int)P[s * d.N + f];
}

template <typename R, int L, int K>
struct Units {
    int a[K], s[K];  // alpha and symmetry index per unit, a < 0: padding
    cplx<R> th[K], tn[K];
    // FP32 (K <= 4; more units would cost occupancy): cosh and
    // sinh of the last ratio12 flip set's delta per unit and the unit's
    // ratio factor cosh(theta + delta) / cosh(theta), kept for accept12's
    // tanh update
    static constexpr bool kMob = sizeof(R) == 4 && K <= 4;
    // FP32 with more units per lane: the same Mobius step, its cosh / sinh
    // factors reloaded from the tables at the (less frequent) accepted move
    // instead of kept in registers from the ratio
    static constexpr bool kMobR = sizeof(R) == 4 && !kMob;
    cplx<R> kc[kMob ? K : 1], ks[kMob ? K : 1], kr[kMob ? K : 1];
    // K <= 2, L <= 32 (the Pfaffian shapes have L = 64 and no registers to
    // spare): W entries of the last ratio12 flip set, loaded with its C
    // and S entries for accept12
    static constexpr bool kPreW = K <= 2 && L <= 32;
    cplx<R> kw0[kPreW ? K : 1], kw1[kPreW ? K : 1];

    __device__ void init(const Dims& d, int ln) {
#pragma unroll
        for (int u = 0; u < K; u++) {
            int h = ln + u * L;
            a[u] = h < d.H ? h % d.n_alpha : -1;
            s[u] = h < d.H ? h / d.n_alpha : 0;
        }
    }
    // theta = b + sum_f sigma_f W_{perm_s(f), alpha} (rbm_base.cpp:161-172,
    // rbm_symmetry.cpp:48-61), then tanh
    __device__ void from_state(const Dims& d, const Lds<R>& l, const cplx<R>* hb, const uint64_t* st) {
#pragma unroll
        for (int u = 0; u < K; u++) {
            if (a[u] < 0) continue;
            cplx<R> acc = hb[a[u]];
            for (int f = 0; f < d.N; f++) {
                int wi = kidx(d, l.P, s[u], f) * d.n_alpha + a[u];
                cplx<R> w = d.w_lds ? l.Wl[wi] : l.Wg[wi];
                if (bitof(st, f)) acc = acc - w;
                else acc = acc + w;
            }
            th[u] = acc;
            tn[u] = ctanh(acc);
        }
    }
    __device__ void ratio12_units(const Dims& d, const Lds<R>& l, int f0, R s0, int f1, R s1, cplx<R>& p,
                                  int& pe) {
        p = {1, 0};
        pe = 0;
#pragma unroll
        for (int u = 0; u < K; u++) {
            if (a[u] < 0) continue;
            int i0 = kidx(d, l.P, s[u], f0) * d.n_alpha + a[u];
            cplx<R> ch = l.C[i0];
            cplx<R> sh = (-s0) * l.S[i0];
            if (f1 >= 0) {
                int i1 = kidx(d, l.P, s[u], f1) * d.n_alpha + a[u];
                cplx<R> c1 = l.C[i1];
                cplx<R> sn1 = (-s1) * l.S[i1];
                cplx<R> nch = ch * c1 + sh * sn1;
                sh = sh * c1 + ch * sn1;
                ch = nch;
            }
            const cplx<R> r = ch + tn[u] * sh;
            if constexpr (kMob) {
                kc[u] = ch;
                ks[u] = sh;
                kr[u] = r;
            }
            p = p * r;
            if (sizeof(R) == 4 && (u & 3) == 3) renorm(p, pe);
        }
    }
    // table rows of site f for the lane's units (symmetric RBM: row of the
    // permuted site; one uniform branch, the loads of all units together)
    __device__ void rows(const Dims& d, const Lds<R>& l, int f, int (&r)[K]) const {
        if (d.n_symm == 1) {
#pragma unroll
            for (int u = 0; u < K; u++) r[u] = f;
        } else {
#pragma unroll
            for (int u = 0; u < K; u++) r[u] = (int)l.P[s[u] * d.N + f];
        }
    }
    // lane-partial product of cosh(theta')/cosh(theta) for one or two flips.
    // K <= 2: every unit's table entries are loaded before any is used
    // (padded units, a < 0, read the entries of alpha 0 and are skipped), so
    // one wait covers them; wider shapes load unit by unit (fewer live
    // registers: the batched form would cost them a wave per SIMD)
    __device__ void ratio12(const Dims& d, const Lds<R>& l, int f0, R s0, int f1, R s1, cplx<R>& p, int& pe) {
        if constexpr (K > 2) {
            ratio12_units(d, l, f0, s0, f1, s1, p, pe);
            return;
        }
        int r0[K], r1[K];
        rows(d, l, f0, r0);
        if (f1 >= 0) rows(d, l, f1, r1);
        cplx<R> c0[K], q0[K], c1[K], q1[K];
        const cplx<R>* W = d.w_lds ? l.Wl : l.Wg;
#pragma unroll
        for (int u = 0; u < K; u++) {
            const int i0 = r0[u] * d.n_alpha + max(a[u], 0);
            c0[u] = l.C[i0];
            q0[u] = l.S[i0];
            if constexpr (kPreW) kw0[u] = W[i0];
            if (f1 >= 0) {
                const int i1 = r1[u] * d.n_alpha + max(a[u], 0);
                c1[u] = l.C[i1];
                q1[u] = l.S[i1];
                if constexpr (kPreW) kw1[u] = W[i1];
            }
        }
        p = {1, 0};
        pe = 0;
#pragma unroll
        for (int u = 0; u < K; u++) {
            if (a[u] < 0) continue;
            cplx<R> ch = c0[u];
            cplx<R> sh = (-s0) * q0[u];
            if (f1 >= 0) {
                cplx<R> sn1 = (-s1) * q1[u];
                cplx<R> nch = ch * c1[u] + sh * sn1;
                sh = sh * c1[u] + ch * sn1;
                ch = nch;
            }
            const cplx<R> r = ch + tn[u] * sh;
            if constexpr (kMob) {
                kc[u] = ch;
                ks[u] = sh;
                kr[u] = r;
            }
            p = p * r;
            if (sizeof(R) == 4 && (u & 3) == 3) renorm(p, pe);
        }
    }
    // FP32 (kMob) after a sweep: tanh from theta again (the accepted moves
    // update tanh incrementally, see accept12_w)
    __device__ void resync() {
        if constexpr (kMob || kMobR) {
#pragma unroll
            for (int u = 0; u < K; u++)
                if (a[u] >= 0) tn[u] = ctanh(th[u]);
        }
    }
    // generic flip set: sites fl[i] for i < 8 with bit i of mask set
    __device__ void ratio_mask(const Dims& d, const Lds<R>& l, const int* fl, const R* sg, int mask,
                               cplx<R>& p, int& pe) const {
        p = {1, 0};
        pe = 0;
#pragma unroll
        for (int u = 0; u < K; u++) {
            if (a[u] < 0) continue;
            cplx<R> ch{1, 0}, sh{0, 0};
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (!((mask >> i) & 1)) continue;
                int ii = kidx(d, l.P, s[u], fl[i]) * d.n_alpha + a[u];
                cplx<R> c = l.C[ii];
                cplx<R> sn = (-sg[i]) * l.S[ii];
                cplx<R> nch = ch * c + sh * sn;
                sh = sh * c + ch * sn;
                ch = nch;
            }
            p = p * (ch + tn[u] * sh);
            if (sizeof(R) == 4 && (u & 3) == 3) renorm(p, pe);
        }
    }
    // accepted move of the flip set (fl[i], sg[i]) for mask bits i < 8
    __device__ void accept_mask(const Dims& d, const Lds<R>& l, const int* fl, const R* sg, int mask) {
        const cplx<R>* W = d.w_lds ? l.Wl : l.Wg;
#pragma unroll
        for (int u = 0; u < K; u++) {
            if (a[u] < 0) continue;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (!((mask >> i) & 1)) continue;
                cplx<R> w = W[kidx(d, l.P, s[u], fl[i]) * d.n_alpha + a[u]];
                th[u] = th[u] - (2 * sg[i]) * w;
            }
            tn[u] = ctanh(th[u]);
        }
    }
    // accepted move: theta -= 2 sigma_f W (rbm_base.cpp:174-192)
    __device__ void accept12(const Dims& d, const Lds<R>& l, int f0, R s0, int f1, R s1) {
        if (d.w_lds)
            accept12_w(d, l, l.Wl, f0, s0, f1, s1);
        else
            accept12_w(d, l, l.Wg, f0, s0, f1, s1);
    }
    // FP64 recomputes tanh(theta) (the chains follow the oracle's state for
    // state); FP32 (kMob) updates it from the ratio12 of the same flip set:
    // tanh(theta + delta) = (t cosh delta + sinh delta) / (cosh delta +
    // t sinh delta), whose denominator is the unit's ratio factor (a move
    // is accepted with probability |prod of factors|^2, so a vanishing
    // factor is almost never divided by); resync() restores tanh(theta)
    // after every sweep, so rounding builds up over one sweep at most
    __device__ void accept12_w(const Dims& d, const Lds<R>& l, const cplx<R>* W, int f0, R s0, int f1, R s1) {
        if constexpr (!kPreW) {  // unit by unit
#pragma unroll
            for (int u = 0; u < K; u++) {
                if (a[u] < 0) continue;
                const int i0 = kidx(d, l.P, s[u], f0) * d.n_alpha + a[u];
                cplx<R> w0 = W[i0];
                th[u] = th[u] - (2 * s0) * w0;
                cplx<R> ch, sh;
                if constexpr (kMobR) {
                    ch = l.C[i0];
                    sh = (-s0) * l.S[i0];
                }
                if (f1 >= 0) {
                    const int i1 = kidx(d, l.P, s[u], f1) * d.n_alpha + a[u];
                    cplx<R> w1 = W[i1];
                    th[u] = th[u] - (2 * s1) * w1;
                    if constexpr (kMobR) {
                        const cplx<R> c1 = l.C[i1], sn1 = (-s1) * l.S[i1];
                        const cplx<R> nch = ch * c1 + sh * sn1;
                        sh = sh * c1 + ch * sn1;
                        ch = nch;
                    }
                }
                if constexpr (kMobR) {
                    // tanh(theta + delta) = (t cosh delta + sinh delta) / (cosh delta + t sinh delta)
                    const cplx<R> num = tn[u] * ch + sh, r = ch + tn[u] * sh;
                    const R inv = __builtin_amdgcn_rcpf(r.re * r.re + r.im * r.im);
                    tn[u] = {(num.re * r.re + num.im * r.im) * inv, (num.im * r.re - num.re * r.im) * inv};
                } else {
                    update_tanh(u);
                }
            }
        } else {  // the W entries came with ratio12 of the same flip set
#pragma unroll
            for (int u = 0; u < K; u++) {
                if (a[u] < 0) continue;
                th[u] = th[u] - (2 * s0) * kw0[u];
                if (f1 >= 0) th[u] = th[u] - (2 * s1) * kw1[u];
                update_tanh(u);
            }
        }
    }
    __device__ void update_tanh(int u) {
        if constexpr (kMob) {
            const cplx<R> num = tn[u] * kc[u] + ks[u], r = kr[u];
            const R inv = __builtin_amdgcn_rcpf(r.re * r.re + r.im * r.im);
            tn[u] = {(num.re * r.re + num.im * r.im) * inv, (num.im * r.re - num.re * r.im) * inv};
        } else {
            tn[u] = ctanh(th[u]);
        }
    }
};

// psi(sigma')/psi(sigma) for the sites of a term selected by mask: the
// visible-bias factor exp(-2 sum sigma_f c_f) (rbm_base.cpp:306-318) times
// the group product of the per-unit cosh ratios
template <typename R, int L, int K, bool PF>
__device__ inline cplx<double> term_ratio(const Dims& d, const Lds<R>& l, const Units<R, L, K>& U,
                                          const cplx<R> (&pv)[16], const uint64_t* st, int ln, const int* fl,
                                          const R* sg, int mask) {
    cplx<R> bias{0, 0};
#pragma unroll
    for (int i = 0; i < 8; i++)
        if ((mask >> i) & 1) bias = bias + (-2 * sg[i]) * l.rb[fl[i]];
    cplx<R> p;
    int pe;
    U.ratio_mask(d, l, fl, sg, mask, p, pe);
    group_cprod<L>(p, pe);
    cplx<R> r = cexp(bias) * p;
    double sc = ::ldexp(1.0, pe);
    cplx<double> out{(double)r.re * sc, (double)r.im * sc};
    if constexpr (PF) {  // times Pf(A(sigma')) / Pf(A(sigma)) (pfaffian.hpp:73-86)
        PfProp<R, false> pp;
        cplx<R> q = pf_ratio<R, false>(l.pf, pv, st, fl, mask, ln, pp);
        out = out * cplx<double>{(double)q.re, (double)q.im};
    }
    return out;
}

// local observables (local_op.cpp:31-60, summed per group as
// local_op_chain.cpp:54-60); lane 0 writes out[2*g], out[2*g+1]
template <typename R, int L, int K, bool PF, bool TL>
__device__ inline void eval_obs_t(const Dims& d, const Lds<R>& l, const Units<R, L, K>& U,
                                  const cplx<R> (&pv)[16], const uint64_t* st, int ln, double* out) {
    const Terms& T = TL ? l.Tl : l.Tg;
    cplx<double> acc{0, 0};
    int g = T.n_terms > 0 ? T.group[0] : 0;
    for (int ti = 0; ti < T.n_terms; ti++) {
        int gt = T.group[ti];
        if (gt != g) {
            if (ln == 0) {
                out[2 * g] = acc.re;
                out[2 * g + 1] = acc.im;
            }
            acc = {0, 0};
            g = gt;
        }
        int k = T.k[ti];
        int fl[8];
        R sg[8];
        int loc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            fl[i] = i < k ? (int)T.sites[ti * 8 + i] : 0;
            int b = i < k ? bitof(st, fl[i]) : 0;
            sg[i] = b ? (R)-1 : (R)1;
            loc |= b << i;
        }
        int base = T.rowbase[ti] + loc;
        int e0 = T.rowptr[base], e1 = T.rowptr[base + 1];
        for (int e = e0; e < e1; e++) {
            int mask = T.mask[e];
            cplx<double> v{T.val[2 * e], T.val[2 * e + 1]};
            if (mask == 0) {
                acc = acc + v;
            } else {
                acc = acc + v * term_ratio<R, L, K, PF>(d, l, U, pv, st, ln, fl, sg, mask);
            }
        }
    }
    if (ln == 0 && T.n_terms > 0) {
        out[2 * g] = acc.re;
        out[2 * g + 1] = acc.im;
    }
}

// FP32 local observables (no Pfaffian factor), same terms and sums as
// eval_obs_t, arranged for latency: the lanes first resolve the terms in
// parallel (lane ln: terms ln + j L, j < kObsWalk -- sites, local
// configuration, entry range), then the terms are taken in order in
// batches of up to kObsBatch terms of one group, each reading its owner
// lane's result by a lane shuffle; the lane-partial cosh ratios of a
// batch's off-diagonal entries are reduced over the group together
// (interleaved butterflies) and multiplied by the visible-bias factor
// exp(-2 sum sigma_f c_f), computed with v_exp_f32 and sincos_f.  A term
// with more than one off-diagonal entry finishes the others one by one.
// Used for K <= 2 units per lane and term tables in LDS (basic and
// symmetric RBM).
constexpr int kObsWalk = 4, kObsBatch = 4;

template <int L, int B, int O = 1>
__device__ inline void group_cprod_batch(cplx<float> (&m)[B], int (&e)[B]) {
    if constexpr (O < L) {
#pragma unroll
        for (int b = 0; b < B; b++) {
            cplx<float> q{partner<O>(m[b].re), partner<O>(m[b].im)};
            int qe = partner<O>(e[b]);
            m[b] = m[b] * q;
            e[b] += qe;
        }
        group_cprod_batch<L, B, 2 * O>(m, e);
    }
}

template <int L, int K, bool TL, bool SYM1>
__device__ inline void eval_obs_batch(const Dims& d, const Lds<float>& l, const Units<float, L, K>& U,
                                      const cplx<float> (&pv)[16], const uint64_t* st, int ln, double* out) {
    const Terms& T = TL ? l.Tl : l.Tg;
    constexpr int J = kObsWalk, B = kObsBatch;
    const int nt = T.n_terms;
    // ---- walk: site bytes (packed), entry range of terms ln + j L
    uint32_t ws0[J], ws1[J];
    int we0[J], we1[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int ti = ln + j * L;
        ws0[j] = ws1[j] = 0;
        we0[j] = we1[j] = 0;
        if (ti < nt) {
            const int k = T.k[ti];
            int loc = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (i < k) {
                    const uint32_t f = T.sites[ti * 8 + i];
                    if (i < 4) ws0[j] |= f << (8 * i);
                    else ws1[j] |= f << (8 * (i - 4));
                    loc |= bitof(st, (int)f) << i;
                }
            }
            const int base = T.rowbase[ti] + loc;
            we0[j] = T.rowptr[base];
            we1[j] = T.rowptr[base + 1];
        }
    }
    // ---- terms in order, by batches of one group
    struct Slot {
        cplx<float> p, bias;  // lane-partial ratio (mantissa p 2^pe), visible-bias exponent
        int pe;
        cplx<double> v;  // matrix element
        bool used;
    };
    cplx<double> acc{0, 0};
    int ti = 0;
    int g = nt > 0 ? T.group[0] : 0;
    // next term of group g into slot S (diagonal entries go straight to acc)
    auto fill = [&](Slot& S) __attribute__((always_inline)) {
        S.p = {1.f, 0.f};
        S.pe = 0;
        S.bias = {0.f, 0.f};
        S.v = {0, 0};
        S.used = false;
        if (ti >= nt || T.group[ti] != g) return;
        const int t = ti++, j = t / L, src = t % L;
        uint32_t s0 = ws0[0], s1 = ws1[0];
        int e0 = we0[0], e1 = we1[0];
#pragma unroll
        for (int jj = 1; jj < J; jj++)
            if (j == jj) {
                s0 = ws0[jj];
                s1 = ws1[jj];
                e0 = we0[jj];
                e1 = we1[jj];
            }
        s0 = (uint32_t)__shfl((int)s0, src, L);
        s1 = (uint32_t)__shfl((int)s1, src, L);
        e0 = __shfl(e0, src, L);
        e1 = __shfl(e1, src, L);
        for (int e = e0; e < e1; e++) {
            const int mask = T.mask[e];
            const cplx<double> v{T.val[2 * e], T.val[2 * e + 1]};
            if (mask == 0) {
                acc = acc + v;
                continue;
            }
            // lane-partial product of cosh(theta')/cosh(theta) over the
            // flipped sites (set bits of mask, in order; the loop over the
            // bits keeps the table loads to the sites that are flipped)
            cplx<float> ch[K], sh[K];
#pragma unroll
            for (int u = 0; u < K; u++) {
                ch[u] = {1.f, 0.f};
                sh[u] = {0.f, 0.f};
            }
            cplx<float> bias{0.f, 0.f};
            for (int m = mask; m; m &= m - 1) {
                const int i = __builtin_ctz(m);
                const int f = (int)(((i < 4 ? s0 : s1) >> (8 * (i & 3))) & 0xffu);
                const float sgn = bitof(st, f) ? -1.f : 1.f;
                // all loads of the site at once (one wait): its bias and the
                // units' table rows (basic RBM: row f; padded units read the
                // row of alpha 0 and are not used)
                const cplx<float> rbv = l.rb[f];
                cplx<float> c[K], sn[K];
#pragma unroll
                for (int u = 0; u < K; u++) {
                    const int row = SYM1 ? f : (int)l.P[U.s[u] * d.N + f];
                    const int ii = row * d.n_alpha + max(U.a[u], 0);
                    c[u] = l.C[ii];
                    sn[u] = l.S[ii];
                }
                bias = bias + (-2 * sgn) * rbv;
#pragma unroll
                for (int u = 0; u < K; u++) {
                    if (U.a[u] < 0) continue;
                    const cplx<float> snu = (-sgn) * sn[u];
                    const cplx<float> nch = ch[u] * c[u] + sh[u] * snu;
                    sh[u] = sh[u] * c[u] + ch[u] * snu;
                    ch[u] = nch;
                }
            }
            cplx<float> p{1.f, 0.f};
            int pe = 0;
#pragma unroll
            for (int u = 0; u < K; u++) {
                if (U.a[u] < 0) continue;
                p = p * (ch[u] + U.tn[u] * sh[u]);
                if ((u & 3) == 3) renorm(p, pe);
            }
            if (!S.used) {
                S.p = p;
                S.pe = pe;
                S.bias = bias;
                S.v = v;
                S.used = true;
            } else {  // a second off-diagonal entry of the term
                group_cprod<L>(p, pe);
                float sn, cs;
                sincos_f(bias.im, sn, cs);
                const float ex = __expf(bias.re);
                const cplx<float> r = cplx<float>{ex * cs, ex * sn} * p;
                const double sc = ::ldexp(1.0, pe);
                acc = acc + v * cplx<double>{(double)r.re * sc, (double)r.im * sc};
            }
        }
    };
    static_assert(B == 4, "four slots per batch");
    while (ti < nt) {
        Slot s0, s1, s2, s3;
        fill(s0);
        fill(s1);
        fill(s2);
        fill(s3);
        cplx<float> bp[B] = {s0.p, s1.p, s2.p, s3.p};
        int bpe[B] = {s0.pe, s1.pe, s2.pe, s3.pe};
#pragma unroll
        for (int b = 0; b < B; b++) renorm(bp[b], bpe[b]);
        group_cprod_batch<L, B>(bp, bpe);
        auto finish = [&](const Slot& S, cplx<float> p, int pe) __attribute__((always_inline)) {
            if (!S.used) return;
            float sn, cs;
            sincos_f(S.bias.im, sn, cs);
            const float ex = __expf(S.bias.re);
            const cplx<float> r = cplx<float>{ex * cs, ex * sn} * p;
            const double sc = ::ldexp(1.0, pe);
            acc = acc + S.v * cplx<double>{(double)r.re * sc, (double)r.im * sc};
        };
        finish(s0, bp[0], bpe[0]);
        finish(s1, bp[1], bpe[1]);
        finish(s2, bp[2], bpe[2]);
        finish(s3, bp[3], bpe[3]);
        if (ti >= nt || T.group[ti] != g) {
            if (ln == 0) {
                out[2 * g] = acc.re;
                out[2 * g + 1] = acc.im;
            }
            acc = {0, 0};
            if (ti < nt) g = T.group[ti];
        }
    }
}

template <typename R, int L, int K, bool PF>
__device__ inline void eval_obs(const Dims& d, const Lds<R>& l, const Units<R, L, K>& U,
                                const cplx<R> (&pv)[16], const uint64_t* st, int ln, double* out) {
    // (with more units per lane the batch would not fit in the registers;
    // term tables in LDS; the basic RBM reads table row f for site f, the
    // symmetric one row P[s][f] of the unit's symmetry image s)
    if constexpr (sizeof(R) == 4 && !PF && K <= 2) {
        if (d.n_terms <= kObsWalk * L && d.t_lds) {
            if (d.n_symm == 1)
                eval_obs_batch<L, K, true, true>(d, l, U, pv, st, ln, out);
            else
                eval_obs_batch<L, K, true, false>(d, l, U, pv, st, ln, out);
            return;
        }
    }
    if (d.t_lds)
        eval_obs_t<R, L, K, PF, true>(d, l, U, pv, st, ln, out);
    else
        eval_obs_t<R, L, K, PF, false>(d, l, U, pv, st, ln, out);
}

// ------------------------------------------------------------ sampler --
// metropolis_sampler.cpp:118-273: one chain per group, random initial
// state, n_warmup + n*step sweeps of N proposal slots; with probability
// bond_flips both ends of a random non-z bond are flipped (taking two
// slots), else one uniform site; accept if u < |psi'/psi|^2.  At kept
// steps the local observables and tanh(theta) of the sample are written.
template <typename R, int L, int K, bool PF>
__global__ void __launch_bounds__(256) k_sample(SampArgs<R> a) {
    extern __shared__ __align__(16) char smem[];
    const Dims d = a.d;
    Lds<R> l = stage<R>(d, a.t, a.T, a.pb, smem);
    int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    int chain = gtid / L, ln = gtid % L;
    if (chain >= a.n_chains) return;
    int64_t spc = a.n_samples / a.n_chains, res = a.n_samples - spc * a.n_chains;
    int64_t nc = spc + (chain == 0 ? res : 0);
    int64_t off = chain == 0 ? 0 : chain * spc + res;
    int64_t total = nc * a.step + a.warmup;
    const uint32_t epoch = a.epoch_dev ? *a.epoch_dev : a.epoch;

    uint64_t st[2] = {0, 0};
    {
        u4 r = philox((uint32_t)chain, epoch, 0xFFFFFFFFu, 0u, a.k0, a.k1);
        uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int q = 0; q < 4; q++) st[q >> 1] |= (uint64_t)w[q] << (32 * (q & 1));
        if (d.N < 64) st[0] &= (((uint64_t)1) << d.N) - 1;
        if (d.N <= 64) st[1] = 0;
        else if (d.N < 128) st[1] &= (((uint64_t)1) << (d.N - 64)) - 1;
    }
    Units<R, L, K> U;
    U.init(d, ln);
    U.from_state(d, l, a.t.hb, st);
    // Pfaffian: inverse of A(sigma) in registers (one chain per wavefront)
    cplx<R> pv[16];
    PfProp<R, true> pp;
    if constexpr (PF) pf_fresh(l.pf, st, ln, pv);

    // The proposals of slots sw0 .. sw0 + L - 1 of a step are drawn ahead,
    // one Philox block per lane (lane ln: slot sw0 + ln), and read by the
    // group with a lane shuffle: code = f0 | (f1 + 1) << 8, f1 = -1 for a
    // single flip, and the acceptance uniform.
    int pcode = 0;
    R pu = 0;
    auto draw = [&](int64_t step, int sw0) {
        u4 r = philox((uint32_t)chain, epoch, (uint32_t)step, (uint32_t)(sw0 + ln), a.k0, a.k1);
        int f0, f1 = -1;
        if ((uint64_t)r.x < a.thr && d.npb > 0) {
            int bi = (int)bounded(r.y, (uint32_t)d.npb);
            f0 = l.pb[2 * bi];
            f1 = l.pb[2 * bi + 1];
        } else {
            f0 = (int)bounded(r.y, (uint32_t)d.N);
        }
        pcode = f0 | ((f1 + 1) << 8);
        pu = uniform<R>(r.z, r.w);
    };

    uint32_t acc = 0, prop = 0;
    int64_t k = 0;
    for (int64_t step = 0; step < total; step++) {
        int sw0 = 0;
        draw(step, 0);
        for (int sw = 0; sw < d.N; sw++) {
            if (sw >= sw0 + L) {
                sw0 = sw & ~(L - 1);
                draw(step, sw0);
            }
            int code = __shfl(pcode, sw - sw0, L);
            R uu = __shfl(pu, sw - sw0, L);
            int f0 = code & 0xFF, f1 = (code >> 8) - 1;
            if (f1 >= 0) sw++;
            prop++;
            R s0 = bitof(st, f0) ? (R)-1 : (R)1;
            R s1 = (f1 >= 0 && bitof(st, f1)) ? (R)-1 : (R)1;
            cplx<R> p;
            int pe;
            U.ratio12(d, l, f0, s0, f1, s1, p, pe);
            // log2 |psi'/psi|^2: group sum of the lanes' log2 |p|^2, plus the
            // visible-bias term 2 Re(-2 sum sigma_f c_f) / ln 2
            R lp = M<R>::log2(norm2(p)) + (R)(2 * pe);
            R lr2 = group_sum<L>(lp);
            R bre = (-2 * s0) * l.rb[f0].re;
            if (f1 >= 0) bre += (-2 * s1) * l.rb[f1].re;
            lr2 += (R)2.8853900817779268 * bre;  // 2 / ln 2
            if constexpr (PF) {
                int fl[2] = {f0, f1};
                cplx<R> q = pf_ratio<R, true>(l.pf, pv, st, fl, f1 >= 0 ? 3 : 1, ln, pp);
                lr2 += M<R>::log2(norm2(q));
            }
            if (uu < M<R>::exp2(lr2)) {
                U.accept12(d, l, f0, s0, f1, s1);
                if constexpr (PF) pf_accept(l.pf, pv, pp, ln);
                st[f0 >> 6] ^= (uint64_t)1 << (f0 & 63);
                if (f1 >= 0) st[f1 >> 6] ^= (uint64_t)1 << (f1 & 63);
                acc++;
                // fresh inverse every pf_refresh accepted moves (metropolis_sampler.cpp:240,
                // abstract_sampler.cpp:103-113)
                if constexpr (PF)
                    if (a.pf_refresh > 0 && acc % (uint32_t)a.pf_refresh == 0) pf_fresh(l.pf, st, ln, pv);
            }
        }
        U.resync();
        if (step >= a.warmup && (step - a.warmup) % a.step == 0) {
            int64_t smp = off + k;
            if (ln == 0)
                for (int w = 0; w < d.nw; w++) a.states[smp * d.nw + w] = st[w];
#pragma unroll
            for (int u = 0; u < K; u++)
                if (U.a[u] >= 0) a.Tout[smp * d.H + ln + u * L] = U.tn[u];
            if (!a.obs_later) eval_obs<R, L, K, PF>(d, l, U, pv, st, ln, a.obs + smp * 2 * a.T.n_groups);
            if constexpr (PF) pf_derivative(l.pf, pv, st, ln, a.Opf + smp * 4 * d.pf_rows);
            k++;
        }
    }
    if (ln == 0) {
        a.acc[chain] = acc;
        a.prop[chain] = prop;
    }
}

// ---------------------------------------------------------- eval states --
// one group per host-given state: theta from scratch, tanh, optionally
// log psi, local observables and the psi ratio for given flips
template <typename R, int L, int K, bool PF>
__global__ void __launch_bounds__(256) k_eval(EvalArgs<R> a) {
    extern __shared__ __align__(16) char smem[];
    const Dims d = a.d;
    Lds<R> l = stage<R>(d, a.t, a.T, a.pb, smem);
    int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t g = gtid / L;
    int ln = gtid % L;
    if (g >= a.n) return;
    uint64_t st[2] = {0, 0};
    for (int w = 0; w < d.nw; w++) st[w] = a.states[g * d.nw + w];
    Units<R, L, K> U;
    U.init(d, ln);
    U.from_state(d, l, a.t.hb, st);
    cplx<R> pv[16];
    if constexpr (PF) pf_fresh(l.pf, st, ln, pv);
    if (a.Tout) {
#pragma unroll
        for (int u = 0; u < K; u++)
            if (U.a[u] >= 0) a.Tout[g * d.H + ln + u * L] = U.tn[u];
    }
    if (a.logpsi) {
        // log psi = sum_f c_f sigma_f + sum_h ln cosh theta_h (rbm_base.cpp:294-304)
        cplx<R> s{0, 0};
#pragma unroll
        for (int u = 0; u < K; u++)
            if (U.a[u] >= 0) s = s + clncosh(U.th[u]);
        for (int f = ln; f < d.N; f += L) {
            R sg = bitof(st, f) ? (R)-1 : (R)1;
            s = s + sg * a.t.lb[f];
        }
        s = group_csum<L>(s);
        cplx<double> lp{(double)s.re, (double)s.im};
        if constexpr (PF) lp = lp + pf_logpf(l.pf, st, ln);  // + ln Pf(A(sigma))
        if (ln == 0) {
            a.logpsi[2 * g] = lp.re;
            a.logpsi[2 * g + 1] = lp.im;
        }
    }
    if constexpr (PF)
        if (a.Opf) pf_derivative(l.pf, pv, st, ln, a.Opf + g * 4 * d.pf_rows);
    if (a.obs) eval_obs<R, L, K, PF>(d, l, U, pv, st, ln, a.obs + g * 2 * a.T.n_groups);
    if (a.ratio) {
        int fl[8];
        R sg[8];
        int mask = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            fl[i] = i < a.n_flips ? a.flips[i] : 0;
            sg[i] = bitof(st, fl[i]) ? (R)-1 : (R)1;
            if (i < a.n_flips) mask |= 1 << i;
        }
        cplx<double> r = term_ratio<R, L, K, PF>(d, l, U, pv, st, ln, fl, sg, mask);
        if (ln == 0) {
            a.ratio[2 * g] = r.re;
            a.ratio[2 * g + 1] = r.im;
        }
    }
}

// ----------------------------------------------------------- O matrix --
// O column of one sample (rbm_base.cpp:194-217, rbm_symmetry.cpp:97-130):
// [ sum_s sigma_{perm_s(basis_i)} | sum_s tanh theta(a,s) |
//   sum_s (P_s sigma)_j tanh theta(a,s) at j + a*N ]
// A block takes batches of kOBatch samples (their tanh rows and states in
// LDS, staged by unrolled register loads; the permutation tables staged
// once per block as bytes, N <= 128); 256 threads loop over the (sample,
// site) items of a batch.  Weight entries:
// item (u, j) runs over the symmetries once per 4 alpha channels, the spin
// (P_s sigma)_j read as a bit of the state words in registers, the 4
// channels' tanh values read unconditionally (no per-channel branch between
// the LDS reads), so the stores of a wave cover consecutive j.
constexpr int kOBatch = 8;
template <typename R>
__host__ __device__ inline size_t build_O_lds(const Dims& d) {
    return (size_t)kOBatch * d.H * sizeof(cplx<R>) + (size_t)kOBatch * d.nw * sizeof(uint64_t) +
           (size_t)d.n_symm * (d.N + d.n_vb);
}

template <typename R>
__device__ inline R spin_of(uint64_t w0, uint64_t w1, int f) {
    return (((f < 64 ? w0 : w1) >> (f & 63)) & 1) ? (R)-1 : (R)1;
}

template <typename R>
__global__ void __launch_bounds__(256) k_build_O(OArgs<R> a) {
    extern __shared__ __align__(16) char smem[];
    const Dims d = a.d;
    const int N = d.N, ns = d.n_symm, nvb = d.n_vb, na = d.n_alpha, H = d.H, nw = d.nw;
    const int nt = blockDim.x;
    cplx<R>* tr = reinterpret_cast<cplx<R>*>(smem);                     // kOBatch x H
    uint64_t* stw = reinterpret_cast<uint64_t*>(tr + kOBatch * H);      // kOBatch x nw
    uint8_t* ip = reinterpret_cast<uint8_t*>(stw + kOBatch * nw);       // ns x N: perm_s^-1(j)
    uint8_t* vb = ip + ns * N;                                          // ns x nvb: perm_s(basis_i)
    {
        constexpr int U = 4;  // loads in flight per thread
        const int ni = ns * N, nv = ns * nvb;
        for (int e0 = threadIdx.x; e0 < ni + nv; e0 += U * nt) {
            int32_t v[U];
#pragma unroll
            for (int i = 0; i < U; i++) {
                const int e = e0 + i * nt;
                v[i] = e < ni ? a.invperm[e] : e < ni + nv ? a.vb_site[e - ni] : 0;
            }
#pragma unroll
            for (int i = 0; i < U; i++)
                if (e0 + i * nt < ni + nv) ip[e0 + i * nt] = (uint8_t)v[i];  // vb follows ip
        }
    }
    const int nb = nvb + na, tail = a.ld - a.P_rbm;
    const int64_t nbatch = (a.S + kOBatch - 1) / kOBatch;
    for (int64_t bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
        const int64_t kb = bt * kOBatch;
        const int nsb = a.S - kb < kOBatch ? (int)(a.S - kb) : kOBatch;
        __syncthreads();  // the previous batch's readers are done
        {
            constexpr int U = 4;
            const int nel = nsb * H;
            const cplx<R>* src = a.T + kb * H;
            int e0 = threadIdx.x;
            for (; e0 + (U - 1) * nt < nel; e0 += U * nt) {  // whole groups: U loads in flight
                cplx<R> v[U];
#pragma unroll
                for (int i = 0; i < U; i++) v[i] = src[e0 + i * nt];
#pragma unroll
                for (int i = 0; i < U; i++) tr[e0 + i * nt] = v[i];
            }
            for (; e0 < nel; e0 += nt) tr[e0] = src[e0];
            for (int e = threadIdx.x; e < nsb * nw; e += nt) stw[e] = a.states[kb * nw + e];
        }
        __syncthreads();
        for (int it = threadIdx.x; it < nsb * N; it += nt) {
            const int u = it / N, j = it - u * N;
            const uint64_t w0 = stw[u * nw], w1 = nw > 1 ? stw[u * nw + 1] : 0;
            const cplx<R>* tu = tr + u * H;
            cplx<R>* o = a.O + (kb + u) * a.ld + nb + j;
            int al0 = 0;
            for (; al0 + 4 <= na; al0 += 4) {
                cplx<R> acc[4] = {};
                int s = 0;
                for (; s + 4 <= ns; s += 4) {  // the LDS reads of 4 symmetries issued together
                    int f[4];
                    cplx<R> t[4][4];
#pragma unroll
                    for (int i = 0; i < 4; i++) f[i] = ip[(s + i) * N + j];
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int q = 0; q < 4; q++) t[i][q] = tu[(s + i) * na + al0 + q];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const R sg = spin_of<R>(w0, w1, f[i]);
#pragma unroll
                        for (int q = 0; q < 4; q++) acc[q] = acc[q] + sg * t[i][q];
                    }
                }
                for (; s < ns; s++) {
                    const R sg = spin_of<R>(w0, w1, ip[s * N + j]);
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[q] = acc[q] + sg * tu[s * na + al0 + q];
                }
#pragma unroll
                for (int q = 0; q < 4; q++) o[(int64_t)(al0 + q) * N] = acc[q];
            }
            for (; al0 < na; al0++) {  // channels past the last group of 4
                cplx<R> acc{0, 0};
                for (int s = 0; s < ns; s++) acc = acc + spin_of<R>(w0, w1, ip[s * N + j]) * tu[s * na + al0];
                o[(int64_t)al0 * N] = acc;
            }
        }
        // biases, the Pfaffian block (pfaffian.cpp:399-415, written by the
        // sampler) and the row padding
        for (int it = threadIdx.x; it < nsb * (nb + tail); it += nt) {
            const int uu = it / (nb + tail), p = it - uu * (nb + tail);
            const uint64_t w0 = stw[uu * nw], w1 = nw > 1 ? stw[uu * nw + 1] : 0;
            const cplx<R>* tu = tr + uu * H;
            const int64_t k = kb + uu;
            cplx<R> v{0, 0};
            int col = p;
            if (p < nvb) {
                R acc = 0;
                for (int s = 0; s < ns; s++) acc += spin_of<R>(w0, w1, vb[s * nvb + p]);
                v = {acc, 0};
            } else if (p < nb) {
                for (int s = 0; s < ns; s++) v = v + tu[p - nvb + s * na];
            } else {
                col = a.P_rbm + (p - nb);
                if (col < a.P && a.Opf) v = a.Opf[k * (a.P - a.P_rbm) + col - a.P_rbm];
            }
            a.O[k * a.ld + col] = v;
        }
    }
}

// ------------------------------------------------------------ moments --
// per-block partial sums over a chunk of samples: sum O, sum |O|^2,
// sum E conj(O) (accumulated in double)
template <typename R>
__global__ void __launch_bounds__(256) k_o_partial(MomArgs<R> a) {
    // block (chunk b = blockIdx.x, parameters 64 blockIdx.y + lane): sums of
    // O - K, |O - K|^2 with the shift K = the chunk's first O (shifted-data
    // variance: no cancellation when O is nearly constant, e.g. saturated
    // hidden units) and of E conj(O), each sample weighted by w (1 without
    // weights).  Wave w takes samples k0 + w, k0 + w + 4, ...; the 4 waves
    // add up in LDS.  part row per (b, p): [K re, K im, sum w d re,
    // sum w d im, sum w |d|^2, sum w E conj(O) re, im, sum w, sum w^2];
    // k_o_final merges the chunks.
    __shared__ double red[3][7][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = (int)blockIdx.y * 64 + lane;
    const bool on = p < a.P;
    const int64_t k0 = (int64_t)blockIdx.x * a.chunk, k1 = k0 + a.chunk < a.S ? k0 + a.chunk : a.S;
    const int pp = on ? p : 0;
    const cplx<R> o0 = a.O[k0 * a.ld + pp];
    const double kr = o0.re, ki = o0.im;
    double s0 = 0, s1 = 0, s2 = 0, e0 = 0, e1 = 0, sw = 0, sw2 = 0;
    const int64_t eo = 2 * (int64_t)a.n_groups;
    auto add = [&](int64_t k, cplx<R> o) {
        const double ore = o.re, oim = o.im;
        const double ere = a.obs[k * eo], eim = a.obs[k * eo + 1];
        const double w = a.w ? a.w[k] : 1.0;
        const double dr = ore - kr, di = oim - ki;
        s0 += w * dr;
        s1 += w * di;
        s2 += w * (dr * dr + di * di);
        e0 += w * (ere * ore + eim * oim);  // E * conj(O)
        e1 += w * (eim * ore - ere * oim);
        sw += w;
        sw2 += w * w;
    };
    int64_t k = k0 + wv;
    for (; k + 12 < k1; k += 16) {  // four rows in flight per wave
        const cplx<R> o_0 = a.O[k * a.ld + pp], o_1 = a.O[(k + 4) * a.ld + pp];
        const cplx<R> o_2 = a.O[(k + 8) * a.ld + pp], o_3 = a.O[(k + 12) * a.ld + pp];
        add(k, o_0);
        add(k + 4, o_1);
        add(k + 8, o_2);
        add(k + 12, o_3);
    }
    for (; k < k1; k += 4) add(k, a.O[k * a.ld + pp]);
    if (wv > 0) {
        double* r = &red[wv - 1][0][lane];
        r[0] = s0;
        r[64] = s1;
        r[128] = s2;
        r[192] = e0;
        r[256] = e1;
        r[320] = sw;
        r[384] = sw2;
    }
    __syncthreads();
    if (wv == 0 && on) {
#pragma unroll
        for (int q = 0; q < 3; q++) {  // fixed order: waves 1, 2, 3
            s0 += red[q][0][lane];
            s1 += red[q][1][lane];
            s2 += red[q][2][lane];
            e0 += red[q][3][lane];
            e1 += red[q][4][lane];
            sw += red[q][5][lane];
            sw2 += red[q][6][lane];
        }
        double* o = a.part + ((int64_t)blockIdx.x * a.P + p) * kOPartRow;
        o[0] = kr;
        o[1] = ki;
        o[2] = s0;
        o[3] = s1;
        o[4] = s2;
        o[5] = e0;
        o[6] = e1;
        o[7] = sw;
        o[8] = sw2;
    }
}

// ------------------------------------------------------------------ S.x --
// One pass over O (rows of ld complex, ld = P rounded up to even, so every
// pair of entries is one 16-byte load in FP32).  Each wavefront streams
// whole rows k: y_k = O_k . x by a wave reduction, then conj(O_k) y_k is
// accumulated in lane registers (FP32: the row stays in registers; FP64:
// the row is re-read from cache).  The waves of a block add their partials
// in LDS; k_sx_colsum sums the block partials and finishes the product
// (regularisation and the <O> term).  Short rows (ld <= 512) run one block
// of 16 waves per CU, so the column sums read one partial row per CU.
template <typename R>
__device__ inline void load_pair(const cplx<R>* o, int p, cplx<R>& a, cplx<R>& b);
template <>
__device__ inline void load_pair<float>(const cplx<float>* o, int p, cplx<float>& a, cplx<float>& b) {
    float4 v = *reinterpret_cast<const float4*>(o + p);
    a = {v.x, v.y};
    b = {v.z, v.w};
}
template <>
__device__ inline void load_pair<double>(const cplx<double>* o, int p, cplx<double>& a, cplx<double>& b) {
    double2 v0 = *reinterpret_cast<const double2*>(o + p);
    double2 v1 = *reinterpret_cast<const double2*>(o + p + 1);
    a = {v0.x, v0.y};
    b = {v1.x, v1.y};
}

template <typename R, int ZM, int WAVES>  // ZM pairs per lane: ld <= 128 * ZM; WAVES per block
__global__ void __launch_bounds__(64 * WAVES) k_sx_partial(SxArgs<R> a) {
    extern __shared__ __align__(16) char smem[];
    if (a.status && a.status->done) return;
    cplx<R>* b = reinterpret_cast<cplx<R>*>(smem);
    cplx<R>* zs = b + a.ld;
    for (int p = threadIdx.x; p < a.ld; p += blockDim.x) {
        b[p] = p < a.P ? cplx<R>{(R)a.x[2 * p], (R)a.x[2 * p + 1]} : cplx<R>{0, 0};
        zs[p] = {0, 0};
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    cplx<R> z[2 * ZM];
#pragma unroll
    for (int i = 0; i < 2 * ZM; i++) z[i] = {0, 0};
    for (int64_t k = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv; k < a.S; k += nwaves) {
        const cplx<R>* o = a.O + k * a.ld;
        cplx<R> dot{0, 0};
        if constexpr (sizeof(R) == 4) {
            cplx<R> v[2 * ZM];
#pragma unroll
            for (int i = 0; i < ZM; i++) {
                int p = 2 * lane + 128 * i;
                if (p < a.ld) load_pair<R>(o, p, v[2 * i], v[2 * i + 1]);
                else v[2 * i] = v[2 * i + 1] = {0, 0};
            }
#pragma unroll
            for (int i = 0; i < ZM; i++) {
                int p = 2 * lane + 128 * i;
                if (p < a.ld) dot = dot + v[2 * i] * b[p] + v[2 * i + 1] * b[p + 1];
            }
            dot = group_csum<64>(dot);
            if (a.wsx) dot = (R)a.wsx[k] * dot;  // weighted sample set (exact enumeration)
#pragma unroll
            for (int i = 0; i < 2 * ZM; i++) z[i] = z[i] + conj(v[i]) * dot;
        } else {
#pragma unroll
            for (int i = 0; i < ZM; i++) {
                int p = 2 * lane + 128 * i;
                if (p < a.ld) {
                    cplx<R> v0, v1;
                    load_pair<R>(o, p, v0, v1);
                    dot = dot + v0 * b[p] + v1 * b[p + 1];
                }
            }
            dot = group_csum<64>(dot);
            if (a.wsx) dot = (R)a.wsx[k] * dot;
#pragma unroll
            for (int i = 0; i < ZM; i++) {
                int p = 2 * lane + 128 * i;
                if (p < a.ld) {
                    cplx<R> v0, v1;
                    load_pair<R>(o, p, v0, v1);
                    z[2 * i] = z[2 * i] + conj(v0) * dot;
                    z[2 * i + 1] = z[2 * i + 1] + conj(v1) * dot;
                }
            }
        }
    }
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        if (wv == w) {
#pragma unroll
            for (int i = 0; i < ZM; i++) {
                int p = 2 * lane + 128 * i;
                if (p < a.ld) {
                    zs[p] = zs[p] + z[2 * i];
                    zs[p + 1] = zs[p + 1] + z[2 * i + 1];
                }
            }
        }
        __syncthreads();
    }
    cplx<R>* out = a.part + (int64_t)blockIdx.x * a.ld;
    for (int p = threadIdx.x; p < a.ld; p += blockDim.x) out[p] = zs[p];
}

// sums the block partials of one chunk of blocks, in FP64
template <typename R>
__global__ void __launch_bounds__(256) k_sx_reduce(SxArgs<R> a) {
    if (a.status && a.status->done) return;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    int64_t per = (a.nblocks + gridDim.y - 1) / gridDim.y;
    int64_t b0 = blockIdx.y * per, b1 = b0 + per < a.nblocks ? b0 + per : a.nblocks;
    double zr = 0, zi = 0;
    for (int64_t b = b0; b < b1; b++) {
        cplx<R> v = a.part[b * a.ld + p];
        zr += v.re;
        zi += v.im;
    }
    a.part2[2 * (blockIdx.y * (int64_t)a.P + p)] = zr;
    a.part2[2 * (blockIdx.y * (int64_t)a.P + p) + 1] = zi;
}

// y = S x from the block partials in one pass: block b owns columns
// 64 b .. 64 b + 63 (lane = column, consecutive addresses); its 16 waves
// sum disjoint slices of the nblocks partial rows, combine them in LDS,
// and wave 0 finishes y_c = z_c / S + reg_c x_c - conj(mu_c) (mu^T x)
// (cg_solver.cpp:301-317).  With pdot, each block also writes its partial
// of x^H y for the CG step.
constexpr int kColWaves = 16;
template <typename R>
__global__ void __launch_bounds__(1024) k_sx_colsum(SxArgs<R> a) {
    if (a.status && a.status->done) return;
    __shared__ double zs[kColWaves][kColBlock][2];
    __shared__ double red[2][64 * kColWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = blockIdx.x * kColBlock + lane;
    const int64_t per = (a.nblocks + kColWaves - 1) / kColWaves, b0 = w * per;
    const int64_t b1 = b0 + per < a.nblocks ? b0 + per : a.nblocks;
    double zr = 0, zi = 0;
    if (c < a.P) {
        int64_t b = b0;
        for (; b + 4 <= b1; b += 4) {
            cplx<R> v0 = a.part[b * a.ld + c], v1 = a.part[(b + 1) * a.ld + c];
            cplx<R> v2 = a.part[(b + 2) * a.ld + c], v3 = a.part[(b + 3) * a.ld + c];
            zr += ((double)v0.re + (double)v1.re) + ((double)v2.re + (double)v3.re);
            zi += ((double)v0.im + (double)v1.im) + ((double)v2.im + (double)v3.im);
        }
        for (; b < b1; b++) {
            cplx<R> v = a.part[b * a.ld + c];
            zr += v.re;
            zi += v.im;
        }
    }
    zs[w][lane][0] = zr;
    zs[w][lane][1] = zi;
    // mu^T x (all threads, P is small)
    double mr = 0, mi = 0;
    for (int p = threadIdx.x; p < a.P; p += blockDim.x) {
        double ur = a.mean[2 * p], ui = a.mean[2 * p + 1], xr = a.x[2 * p], xi = a.x[2 * p + 1];
        mr += ur * xr - ui * xi;
        mi += ur * xi + ui * xr;
    }
    mr = group_sum<64>(mr);
    mi = group_sum<64>(mi);
    if (lane == 0) {
        red[0][w] = mr;
        red[1][w] = mi;
    }
    __syncthreads();
    if (w != 0) return;
    mr = 0;
    mi = 0;
    for (int q = 0; q < kColWaves; q++) {
        mr += red[0][q];
        mi += red[1][q];
    }
    double dr = 0, di = 0;
    if (c < a.P) {
        zr = 0;
        zi = 0;
        for (int q = 0; q < kColWaves; q++) {
            zr += zs[q][lane][0];
            zi += zs[q][lane][1];
        }
        double inv = 1.0 / (double)a.S;
        const double r1 = a.reg_dev ? a.reg_dev[0] : a.r1, r2 = a.reg_dev ? a.reg_dev[1] : a.r2,
                     rd = a.reg_dev ? a.reg_dev[2] : a.rd;
        double reg = a.var[c] * (c < a.n_neural ? r1 : r1 + rd) + (*a.maxvar) * r2;
        double xr = a.x[2 * c], xi = a.x[2 * c + 1];
        double ur = a.mean[2 * c], ui = -a.mean[2 * c + 1];  // conj(mu_c)
        double yr = zr * inv + reg * xr - (ur * mr - ui * mi);
        double yi = zi * inv + reg * xi - (ur * mi + ui * mr);
        a.y[2 * c] = yr;
        a.y[2 * c + 1] = yi;
        dr = xr * yr + xi * yi;  // conj(x) y
        di = xr * yi - xi * yr;
    }
    if (a.pdot) {
        dr = group_sum<64>(dr);
        di = group_sum<64>(di);
        if (lane == 0) {
            a.pdot[2 * blockIdx.x] = dr;
            a.pdot[2 * blockIdx.x + 1] = di;
        }
    }
}

}  // namespace rbmgpu
