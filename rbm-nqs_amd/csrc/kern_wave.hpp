// kern_wave.hpp -- FP32 Metropolis sampler with one chain per wavefront
// (metropolis_sampler.cpp:118-273, the same proposals, Philox streams and
// kept-sample outputs as k_sample), built around wave-uniform control.
//
// Lane ln holds the hidden units h = ln + 64 u (u < K), so the table entry
// of unit u for flipped site f is lane-affine:
//   symmetric RBM (64 % alpha == 0): alpha = ln % alpha_n and symmetry
//     s = ln / alpha_n + u (64 / alpha_n), entry perm_s(f) alpha_n + alpha,
//     the permutation byte at perm + (ln / alpha_n) N + f + u (64 / alpha_n) N;
//   basic RBM: entry f H + ln + 64 u.
// The chain's state, the proposal (read from the drawing lane with
// v_readlane) and the acceptance test are wave-uniform scalars, so the
// proposal loop has no divergent branches and no exec-mask bookkeeping.
//
// Acceptance needs only |psi'/psi|^2 = prod_h |cosh(theta_h + d_h) /
// cosh(theta_h)|^2 times the visible-bias factor: each lane multiplies the
// squared moduli of its units, takes log2, and the wave sums the logs by
// DPP (row sums, then row_bcast 15 / 31).
//
// Bond proposals of the symmetric RBM use pair tables: the image of the
// proposal bond b = (f0, f1) under symmetry s is an ordered site pair
// (perm_s f0, perm_s f1); each distinct image pair p and alpha hold
//   combo 0 (s0 == s1): cosh(2(W0 + W1)), sinh(2(W0 + W1)), W0 + W1
//   combo 1 (s0 != s1): cosh(2(W0 - W1)), sinh(2(W0 - W1)), W0 - W1
// (addition theorems of the single-site tables; W0 = W of the pair's lead
// site), so a two-spin move costs one table entry per unit, like a single
// flip: delta = -2 s0 W_combo, cosh ratio = C + tanh(theta) (-s0 S).
#pragma once
#include "kern_impl.hpp"

namespace rbmgpu {

struct WaveLds {
    const float4* CS;   // (C, S) of entry e = site * alpha_n + alpha (basic: site * H + h)
    const float4* PCS;  // pair tables: combo * npair * alpha_n + pair * alpha_n + alpha
    const float2* W;
    const float2* PW;
    const float2* rb;
    const int32_t* pb;
    const uint8_t* P;   // permutations, rows padded to ns_pad
    const uint8_t* PM;  // pair map [s * npb + b], rows padded to ns_pad
    Terms T;            // local-operator tables (LDS)
    const int16_t* tb;  // per term: the proposal bond of its two sites, -1
    const float2* fac;  // per entry: matrix element x exp(-2 sum_f sigma_f c_f) of its flipped sites
};

template <int K>
__device__ inline WaveLds wave_stage(const SampArgs<float>& a, char* smem) {
    const Dims& d = a.d;
    const int ne = wave_n_entries(d), nreal = d.N * d.n_alpha, na = d.n_alpha;
    const int np = a.npair * na * 2;
    const int nsp = wave_ns_pad(d, K);
    char* at = smem;
    WaveLds l;
    float4* CS = reinterpret_cast<float4*>(at);
    at += (size_t)ne * 16;
    float4* PCS = reinterpret_cast<float4*>(at);
    at += (size_t)np * 16;
    float2* W = reinterpret_cast<float2*>(at);
    at += (size_t)ne * 8;
    float2* PW = reinterpret_cast<float2*>(at);
    at += (size_t)np * 8;
    const cplx<float>* C = a.t.C;
    const cplx<float>* S = a.t.S;
    const cplx<float>* Wg = a.t.W;
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        if (i < nreal) {
            CS[i] = make_float4(C[i].re, C[i].im, S[i].re, S[i].im);
            W[i] = make_float2(Wg[i].re, Wg[i].im);
        } else {
            CS[i] = make_float4(1.f, 0.f, 0.f, 0.f);
            W[i] = make_float2(0.f, 0.f);
        }
    }
    for (int i = threadIdx.x; i < a.npair * na; i += blockDim.x) {
        const int p = i / na, al = i - p * na;
        const int e0 = a.pairs[2 * p] * na + al, e1 = a.pairs[2 * p + 1] * na + al;
        const cplx<float> c0 = C[e0], s0 = S[e0], c1 = C[e1], s1 = S[e1];
        const cplx<float> cc = c0 * c1, ss = s0 * s1, sc = s0 * c1, cs = c0 * s1;
        const int i1 = a.npair * na + i;  // combo 1 half
        PCS[i] = make_float4(cc.re + ss.re, cc.im + ss.im, sc.re + cs.re, sc.im + cs.im);
        PCS[i1] = make_float4(cc.re - ss.re, cc.im - ss.im, sc.re - cs.re, sc.im - cs.im);
        PW[i] = make_float2(Wg[e0].re + Wg[e1].re, Wg[e0].im + Wg[e1].im);
        PW[i1] = make_float2(Wg[e0].re - Wg[e1].re, Wg[e0].im - Wg[e1].im);
    }
    l.CS = CS;
    l.PCS = PCS;
    l.W = W;
    l.PW = PW;
    {
        float2* rb = reinterpret_cast<float2*>(at);
        for (int i = threadIdx.x; i < d.N; i += blockDim.x) rb[i] = make_float2(a.t.rb[i].re, a.t.rb[i].im);
        at += (size_t)d.N * 8;
        l.rb = rb;
    }
    const Terms& T = a.T;
    l.T = T;
    {
        // the entry factors v_e exp(-2 sum_{f flipped} sigma_f c_f): the row
        // (term t, local configuration loc) fixes the signs of its sites
        float2* fac = reinterpret_cast<float2*>(at);
        for (int r = threadIdx.x; r < d.n_terms * 256; r += blockDim.x) {
            const int t = r >> 8, loc = r & 255, k = T.k[t];
            if (loc >> k) continue;
            const int base = T.rowbase[t] + loc;
            for (int e = T.rowptr[base]; e < T.rowptr[base + 1]; e++) {
                const int mask = T.mask[e];
                float br = 0.f, bi = 0.f;
                for (int i = 0; i < k; i++)
                    if ((mask >> i) & 1) {
                        const float m2 = ((loc >> i) & 1) ? 2.f : -2.f;
                        const cplx<float> c = a.t.rb[T.sites[t * 8 + i]];
                        br += m2 * c.re;
                        bi += m2 * c.im;
                    }
                float sn, cs;
                sincos_f(bi, sn, cs);
                const float ex = __expf(br);
                const cplx<float> f = cplx<float>{(float)T.val[2 * e], (float)T.val[2 * e + 1]} *
                                      cplx<float>{ex * cs, ex * sn};
                fac[e] = make_float2(f.re, f.im);
            }
        }
        at += (size_t)d.n_ent * 8;
        l.fac = fac;
    }
    l.T.val = lds_copy(at, T.val, 2 * (size_t)d.n_ent);
    l.T.rowptr = lds_copy(at, T.rowptr, d.n_rowptr);
    l.T.k = lds_copy(at, T.k, d.n_terms);
    l.T.group = lds_copy(at, T.group, d.n_terms);
    l.T.rowbase = lds_copy(at, T.rowbase, d.n_terms);
    l.pb = lds_copy(at, a.pb, 2 * (size_t)d.npb);
    l.tb = lds_copy(at, a.tbond, d.n_terms);
    l.T.sites = lds_copy(at, T.sites, 8 * (size_t)d.n_terms);
    l.T.mask = lds_copy(at, T.mask, d.n_ent);
    l.P = l.PM = reinterpret_cast<const uint8_t*>(at);
    if (d.n_symm > 1) {
        uint8_t* P = reinterpret_cast<uint8_t*>(at);
        for (int i = threadIdx.x; i < nsp * d.N; i += blockDim.x) {
            const int sy = i / d.N;
            P[i] = a.t.perm[(sy < d.n_symm ? i : i - sy * d.N)];
        }
        at += (size_t)nsp * d.N;
        l.P = P;
        if (a.npair) {
            uint8_t* PM = reinterpret_cast<uint8_t*>(at);
            for (int i = threadIdx.x; i < nsp * d.npb; i += blockDim.x) {
                const int sy = i / d.npb;
                PM[i] = a.pmap[(sy < d.n_symm ? i : i - sy * d.npb)];
            }
            at += (size_t)nsp * d.npb;
            l.PM = PM;
        }
    }
    __syncthreads();
    return l;
}

__device__ inline int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
// LDS byte address of an LDS pointer, and loads at LDS byte addresses
__device__ inline uint32_t lds_addr(const void* p) {
    return (uint32_t)(size_t)(__attribute__((address_space(3))) const char*)p;
}
typedef float lds_v4f __attribute__((ext_vector_type(4)));
typedef float lds_v2f __attribute__((ext_vector_type(2)));
__device__ inline float4 lds_f4(uint32_t a) {
    const lds_v4f v = *(__attribute__((address_space(3))) const lds_v4f*)(size_t)a;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline float2 lds_f2(uint32_t a) {
    const lds_v2f v = *(__attribute__((address_space(3))) const lds_v2f*)(size_t)a;
    return make_float2(v.x, v.y);
}
// -1 for a set spin bit, +1 otherwise, and v with its sign flipped by a set
// bit, as integer operations (scalar registers for wave-uniform bits)
__device__ inline float sgn_of(int bit) { return __int_as_float(bit ? (int)0xBF800000 : 0x3F800000); }
__device__ inline float flip_if(float v, int bit) { return __int_as_float(__float_as_int(v) ^ (bit << 31)); }
__device__ inline float rfl(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ inline double rfl(double v) {
    int2 h = *reinterpret_cast<int2*>(&v);
    h.x = __builtin_amdgcn_readfirstlane(h.x);
    h.y = __builtin_amdgcn_readfirstlane(h.y);
    return *reinterpret_cast<double*>(&h);
}

// sum over the 64 lanes, returned as a wave-uniform value: DPP quad
// permutes and row mirrors leave every lane of a 16-lane row with the row
// sum, row_bcast 15 / 31 carry the row sums up, lane 63 ends with the total
template <int CTRL, int RMASK>
__device__ inline float dppf(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, RMASK, 0xF, false));
}
__device__ inline float wave_sum(float v) {
    v += dppf<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
    v += dppf<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
    v += dppf<0x141, 0xF>(v);  // row_half_mirror
    v += dppf<0x140, 0xF>(v);  // row_mirror
    v += dppf<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
    v += dppf<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline int wbit(uint64_t s0, uint64_t s1, int f) {
    return (int)(((f < 64 ? s0 : s1) >> (f & 63)) & 1);
}

template <int K, bool SYM>
struct WaveUnits {
    cplx<float> th[K], tn[K];
    cplx<float> rr[K];  // the last proposal's unit factors (the Mobius step's denominators)
    float rn[K];        // their squared moduli
    uint32_t ea[K];     // and the LDS byte addresses of their (C, S) table records
    int pbase, mbase, aoff, npa, ln, lg;  // lg = log2(alpha_n) (symmetric), npa = npair alpha_n
    uint32_t lcs, lpcs;  // LDS byte addresses of this lane's first CS / PCS record (alpha aoff)
    int ustride, mstride;  // symmetric: byte strides of one unit in perm / pair map
    bool vlast;            // the lane's last unit is a real one

    __device__ void init(const Dims& d, const WaveLds& l, int ln_, int npair) {
        ln = ln_;
        npa = 0;
        if constexpr (SYM) {
            const int na = d.n_alpha, g = 64 / na;
            npa = npair * na;
            lg = __builtin_ctz(na);
            aoff = ln & (na - 1);
            const int sl = ln / na;
            pbase = sl * d.N;
            mbase = sl * d.npb;
            ustride = g * d.N;
            mstride = g * d.npb;
            vlast = sl + (K - 1) * g < d.n_symm;
        } else {
            aoff = lg = pbase = mbase = ustride = mstride = 0;
            vlast = ln + 64 * (K - 1) < d.H;
        }
        lcs = lds_addr(l.CS + aoff);
        lpcs = lds_addr(l.PCS + aoff);
    }
    // LDS byte address of the (C, S) record of unit u: single flip of site f
    // (symmetric RBM: one shift-add from the permutation byte), pair-table
    // bond flip b
    __device__ uint32_t entry_a(const Dims& d, const WaveLds& l, int u, int f) const {
        if constexpr (SYM) return ((uint32_t)l.P[pbase + f + u * ustride] << (lg + 4)) + lcs;
        else return lds_addr(l.CS + entry(d, l, u, f));
    }
    __device__ uint32_t pentry_a(const Dims& d, const WaveLds& l, int u, int b, int combo) const {
        return ((uint32_t)l.PM[mbase + b + u * mstride] << (lg + 4)) + (lpcs + (combo ? 16u * npa : 0u));
    }
    // single-site entry of unit u for site f
    __device__ int entry(const Dims& d, const WaveLds& l, int u, int f) const {
        if constexpr (SYM) return ((int)l.P[pbase + f + u * ustride] << lg) + aoff;
        else return f * d.H + ln + 64 * u;
    }
    // pair-table entry of unit u for proposal bond b
    __device__ int pentry(const Dims& d, const WaveLds& l, int u, int b, int combo) const {
        return ((int)l.PM[mbase + b + u * mstride] << lg) + (aoff + (combo ? npa : 0));
    }
    __device__ cplx<float> hbias(const Dims& d, const cplx<float>* hb, int u) const {
        if constexpr (SYM) return hb[aoff];
        else return hb[min(ln + 64 * u, d.H - 1)];
    }
    // theta = b + sum_f sigma_f W (rbm_base.cpp:161-172, rbm_symmetry.cpp:48-61)
    __device__ void from_state(const Dims& d, const WaveLds& l, const cplx<float>* hb, uint64_t s0, uint64_t s1) {
#pragma unroll
        for (int u = 0; u < K; u++) {
            const bool ok = u < K - 1 || vlast;
            cplx<float> acc = ok ? hbias(d, hb, u) : cplx<float>{0.f, 0.f};
            for (int f = 0; f < d.N; f++) {
                const float2 w = l.W[entry(d, l, u, f)];
                if (wbit(s0, s1, f)) acc = {acc.re - w.x, acc.im - w.y};
                else acc = {acc.re + w.x, acc.im + w.y};
            }
            th[u] = acc;
            tn[u] = ctanh(acc);
        }
    }
    __device__ void resync() {
#pragma unroll
        for (int u = 0; u < K; u++) tn[u] = ctanh(th[u]);
    }
    // unit factor cosh(theta + delta) / cosh(theta) = C + tanh(theta) (ns0 S)
    // of a table entry (C, S)
    __device__ cplx<float> factor1(int u, float4 cs, float ns0) const {
        const float qr = tn[u].re * cs.z - tn[u].im * cs.w;
        const float qi = tn[u].re * cs.w + tn[u].im * cs.z;
        return {fmaf(ns0, qr, cs.x), fmaf(ns0, qi, cs.y)};
    }
    // log2 of the lane's prod_u |C + tanh (ns0 S)|^2 for one table entry per
    // unit (single flip, or a pair entry)
    template <class E>
    __device__ float lratio1(E e, float ns0) {
        float lsum = 0.f, prod = 1.f;
#pragma unroll
        for (int u = 0; u < K; u++) {
            ea[u] = e(u);
            const cplx<float> r = factor1(u, lds_f4(ea[u]), ns0);
            rr[u] = r;
            float n = r.re * r.re + r.im * r.im;
            rn[u] = n;
            if (u == K - 1 && !vlast) n = 1.f;
            prod *= n;
            if ((u & 3) == 3 || u == K - 1) {
                lsum += __log2f(prod);
                prod = 1.f;
            }
        }
        return lsum;
    }
    // cosh / sinh of -2 (s0 W_e0 + s1 W_e1) per unit by the addition theorem
    __device__ void compose2(const WaveLds& l, int e0, int e1, float ns0, float ns1, cplx<float>& ch,
                             cplx<float>& sh) const {
        const float4 a = l.CS[e0], b = l.CS[e1];
        const cplx<float> c0{a.x, a.y}, q0{ns0 * a.z, ns0 * a.w}, c1{b.x, b.y}, q1{ns1 * b.z, ns1 * b.w};
        ch = c0 * c1 + q0 * q1;
        sh = q0 * c1 + c0 * q1;
    }
    __device__ float lratio2(const Dims& d, const WaveLds& l, int f0, int f1, float ns0, float ns1) {
        float lsum = 0.f, prod = 1.f;
#pragma unroll
        for (int u = 0; u < K; u++) {
            cplx<float> ch, sh;
            compose2(l, entry(d, l, u, f0), entry(d, l, u, f1), ns0, ns1, ch, sh);
            const cplx<float> r = ch + tn[u] * sh;
            rr[u] = r;
            float n = r.re * r.re + r.im * r.im;
            rn[u] = n;
            if (u == K - 1 && !vlast) n = 1.f;
            prod *= n;
            if ((u & 3) == 3 || u == K - 1) {
                lsum += __log2f(prod);
                prod = 1.f;
            }
        }
        return lsum;
    }
    // accepted move: theta -= 2 s0 W (rbm_base.cpp:174-192) and the Mobius
    // step tanh(theta + delta) = (t ch + sh) / r with r = ch + t sh, the
    // proposal's unit factor kept in rr (resync() restores tanh(theta) after
    // every sweep)
    // (the squared modulus of r kept from the ratio)
    __device__ void mobius_num(int u, cplx<float> num) {
        const cplx<float> r = rr[u];
        const float inv = __builtin_amdgcn_rcpf(rn[u]);
        tn[u] = {(num.re * r.re + num.im * r.im) * inv, (num.im * r.re - num.re * r.im) * inv};
    }
    __device__ void mobius(int u, cplx<float> ch, cplx<float> sh) { mobius_num(u, tn[u] * ch + sh); }
    // accepted move of the records lratio1 kept: single flip or pair-table
    // bond flip, the W record at (CS address) / 2 + wofs (the W tables hold
    // 8-byte records in the CS tables' order)
    __device__ void accept1(uint32_t wofs, float s0) {
        const float ns0 = -s0, t2 = -2.f * s0;
#pragma unroll
        for (int u = 0; u < K; u++) {
            const float4 cs = lds_f4(ea[u]);
            const float2 w = lds_f2((ea[u] >> 1) + wofs);
            th[u] = {fmaf(t2, w.x, th[u].re), fmaf(t2, w.y, th[u].im)};
            // tanh(theta) C + ns0 S
            const cplx<float> t = tn[u];
            mobius_num(u, {fmaf(ns0, cs.z, fmaf(t.re, cs.x, -t.im * cs.y)),
                           fmaf(ns0, cs.w, fmaf(t.re, cs.y, t.im * cs.x))});
        }
    }
    __device__ void accept2(const Dims& d, const WaveLds& l, int f0, int f1, float s0, float s1) {
#pragma unroll
        for (int u = 0; u < K; u++) {
            const int e0 = entry(d, l, u, f0), e1 = entry(d, l, u, f1);
            const float2 w0 = l.W[e0], w1 = l.W[e1];
            th[u] = {th[u].re - 2.f * (s0 * w0.x + s1 * w1.x), th[u].im - 2.f * (s0 * w0.y + s1 * w1.y)};
            cplx<float> ch, sh;
            compose2(l, e0, e1, -s0, -s1, ch, sh);
            mobius(u, ch, sh);
        }
    }
};

// local observables of the chain's current state (local_op.cpp:31-60,
// summed per group as local_op_chain.cpp:54-60): terms in order; an
// off-diagonal entry's lane-partial ratio, with the entry's factor (matrix
// element times the visible-bias factor exp(-2 sum sigma_f c_f), a table
// built at staging) folded into lane 0's partial, waits in one of four
// slots that are reduced over the wave together; lane 0 writes out[2 g],
// out[2 g + 1].
//
// Symmetric RBM with pair tables: the lanes first resolve the Hamiltonian's
// terms in parallel (lane ln: terms ln and ln + 64 -- sites, local
// configuration, row): diagonal-only rows are summed there, and a row with
// a single off-diagonal entry that flips a proposal bond gets a code for
// the pair-table path, which then runs term after term with no dependent
// table reads in between (the codes come by v_readlane); the other terms
// take the generic path.
template <int K, bool SYM>
__device__ inline void wave_obs(const Dims& d, const WaveLds& l, const WaveUnits<K, SYM>& U, uint64_t st0,
                                uint64_t st1, int npair, double* out) {
    constexpr int B = 4, J = 2;
    const Terms& T = l.T;
    const int nt = T.n_terms;
    cplx<float> bp[B];
    int bpe[B];
    int nb = 0;
    double ar = 0, ai = 0;
    auto flush = [&]() __attribute__((always_inline)) {
        if (nb == 0) return;
#pragma unroll
        for (int i = 0; i < B; i++)
            if (i >= nb) {
                bp[i] = {0.f, 0.f};
                bpe[i] = 0;
            }
        group_cprod_batch<64, B>(bp, bpe);
#pragma unroll
        for (int i = 0; i < B; i++) {
            const double sc = ::ldexp(1.0, bpe[i]);
            ar += (double)bp[i].re * sc;
            ai += (double)bp[i].im * sc;
        }
        nb = 0;
    };
    auto push = [&](cplx<float> p, int pe, int e) __attribute__((always_inline)) {
        if (U.ln == 0) p = p * cplx<float>{l.fac[e].x, l.fac[e].y};
        renorm(p, pe);
#pragma unroll
        for (int i = 0; i < B; i++)
            if (i == nb) {
                bp[i] = p;
                bpe[i] = pe;
            }
        if (++nb == B) flush();
    };
    // the pair-table product of proposal bond tb (lead spin bit sl)
    auto pair_ratio = [&](int tb, int combo, int sl, cplx<float>& p, int& pe) __attribute__((always_inline)) {
        p = {1.f, 0.f};
        pe = 0;
#pragma unroll
        for (int u = 0; u < K; u++) {
            cplx<float> r = U.factor1(u, l.PCS[U.pentry(d, l, u, tb, combo)], sl ? 1.f : -1.f);
            if (u == K - 1 && !U.vlast) r = {1.f, 0.f};
            p = p * r;
            if ((u & 3) == 3) renorm(p, pe);
        }
    };

    // ---- walk (codes: -1 generic, -2 summed here, >= 0 pair path)
    const bool walk = SYM && npair > 0 && nt <= 64 * J && d.n_ent < (1 << 14);
    int wc[J];
    double dr = 0, di = 0;
#pragma unroll
    for (int j = 0; j < J; j++) {
        wc[j] = -1;
        const int t = U.ln + 64 * j;
        if (walk && t < nt && T.group[t] == 0) {
            const int k = T.k[t];
            int loc = 0;
            for (int i = 0; i < k; i++) loc |= wbit(st0, st1, (int)T.sites[t * 8 + i]) << i;
            const int base = T.rowbase[t] + loc, e0 = T.rowptr[base], e1 = T.rowptr[base + 1];
            bool diag = true;
            for (int e = e0; e < e1; e++) diag = diag && T.mask[e] == 0;
            if (diag) {
                for (int e = e0; e < e1; e++) {
                    dr += T.val[2 * e];
                    di += T.val[2 * e + 1];
                }
                wc[j] = -2;
            } else {
                const int tb = l.tb[t];
                if (tb >= 0 && k == 2 && e1 - e0 == 1 && T.mask[e0] == 3) {
                    const int sl = wbit(st0, st1, l.pb[2 * tb]), so = wbit(st0, st1, l.pb[2 * tb + 1]);
                    wc[j] = tb | ((sl == so ? 0 : 1) << 15) | (sl << 16) | (e0 << 17);
                }
            }
        }
    }
    auto code = [&](int t) __attribute__((always_inline)) {
        return t < 64 ? __builtin_amdgcn_readlane(wc[0], t) : __builtin_amdgcn_readlane(wc[1], t - 64);
    };
    if (walk) {
        ar = group_sum<64>(dr);
        ai = group_sum<64>(di);
        for (int t = 0; t < nt; t++) {
            const int c = code(t);
            if (c < 0) continue;
            cplx<float> p;
            int pe;
            pair_ratio(c & 0x7fff, (c >> 15) & 1, (c >> 16) & 1, p, pe);
            push(p, pe, c >> 17);
        }
        flush();
    }

    // ---- the other terms, in order
    int g = nt > 0 ? rfl(T.group[0]) : 0;
    for (int ti = 0; ti < nt; ti++) {
        const int gt = rfl(T.group[ti]);
        if (gt != g) {
            flush();
            if (U.ln == 0) {
                out[2 * g] = ar;
                out[2 * g + 1] = ai;
            }
            ar = ai = 0;
            g = gt;
        }
        if (walk && code(ti) != -1) continue;
        const int k = rfl(T.k[ti]);
        int fl[8];
        int loc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            fl[i] = i < k ? rfl((int)T.sites[ti * 8 + i]) : 0;
            loc |= (i < k ? wbit(st0, st1, fl[i]) : 0) << i;
        }
        const int base = rfl(T.rowbase[ti]) + loc;
        const int e0 = rfl(T.rowptr[base]), e1 = rfl(T.rowptr[base + 1]);
        const int tb = SYM && npair ? rfl((int)l.tb[ti]) : -1;
        for (int e = e0; e < e1; e++) {
            const int mask = rfl((int)T.mask[e]);
            if (mask == 0) {
                ar += rfl(T.val[2 * e]);
                ai += rfl(T.val[2 * e + 1]);
                continue;
            }
            cplx<float> p{1.f, 0.f};
            int pe = 0;
            if (tb >= 0 && k == 2 && mask == 3) {
                // a proposal bond: one pair-table entry per unit
                const int lead = l.pb[2 * tb];
                const int other = lead == fl[0] ? fl[1] : fl[0];
                const int sl = wbit(st0, st1, lead), so = wbit(st0, st1, other);
                pair_ratio(tb, sl == so ? 0 : 1, sl, p, pe);
            } else {
#pragma unroll
                for (int u = 0; u < K; u++) {
                    cplx<float> ch{1.f, 0.f}, sh{0.f, 0.f};
                    for (int m = mask; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        const float ns = wbit(st0, st1, fl[i]) ? 1.f : -1.f;
                        const float4 cs = l.CS[U.entry(d, l, u, fl[i])];
                        const cplx<float> c{cs.x, cs.y}, q{ns * cs.z, ns * cs.w};
                        const cplx<float> nch = ch * c + sh * q;
                        sh = sh * c + ch * q;
                        ch = nch;
                    }
                    cplx<float> r = ch + U.tn[u] * sh;
                    if (u == K - 1 && !U.vlast) r = {1.f, 0.f};
                    p = p * r;
                    if ((u & 3) == 3) renorm(p, pe);
                }
            }
            push(p, pe, e);
        }
    }
    flush();
    if (U.ln == 0 && nt > 0) {
        out[2 * g] = ar;
        out[2 * g + 1] = ai;
    }
}

// PAIRS (symmetric RBM with pair tables, npair > 0): every proposal's ratio
// and accepted move run through one table entry per unit, so the accept
// path is a single block (old and new th / tn never live side by side,
// which keeps the loop-carried registers in place)
template <int K, bool SYM, bool PAIRS>
__global__ void __launch_bounds__(1024) k_sample_wave(SampArgs<float> a) {
    extern __shared__ __align__(16) char smem[];
    const Dims d = a.d;
    const WaveLds l = wave_stage<K>(a, smem);
    const int chain = rfl((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (chain >= a.n_chains) return;
    const int ln = threadIdx.x & 63;
    const int64_t spc = a.n_samples / a.n_chains, res = a.n_samples - spc * a.n_chains;
    const int64_t nc = spc + (chain == 0 ? res : 0);
    const int64_t off = chain == 0 ? 0 : chain * spc + res;
    const int64_t total = nc * a.step + a.warmup;
    const uint32_t epoch = a.epoch_dev ? *a.epoch_dev : a.epoch;

    uint64_t st0, st1;
    {
        const u4 r = philox((uint32_t)chain, epoch, 0xFFFFFFFFu, 0u, a.k0, a.k1);
        st0 = rfl((int)r.x) & 0xffffffffull;
        st0 |= (uint64_t)(uint32_t)rfl((int)r.y) << 32;
        st1 = (uint64_t)(uint32_t)rfl((int)r.z) | ((uint64_t)(uint32_t)rfl((int)r.w) << 32);
        if (d.N < 64) st0 &= (((uint64_t)1) << d.N) - 1;
        if (d.N <= 64) st1 = 0;
        else if (d.N < 128) st1 &= (((uint64_t)1) << (d.N - 64)) - 1;
    }
    WaveUnits<K, SYM> U;
    U.init(d, l, ln, a.npair);
    // W record address = (CS record address) / 2 + offset, per table
    const uint32_t wofs1 = lds_addr(l.W) - lds_addr(l.CS) / 2, wofs2 = lds_addr(l.PW) - lds_addr(l.PCS) / 2;
    U.from_state(d, l, a.t.hb, st0, st1);
    static_assert(SYM || !PAIRS, "pair tables belong to the symmetric RBM");

    // proposals of slots sw0 .. sw0 + 63 are drawn ahead, one Philox block
    // per lane (lane ln: slot sw0 + ln): code = f0 | (f1 + 1) << 8 | b << 16
    // (f1 = -1: single flip; b: proposal bond) and log2 of the uniform
    int pcode = 0;
    float plu = 0.f;
    auto draw = [&](int64_t step, int sw0) {
        const u4 r = philox((uint32_t)chain, epoch, (uint32_t)step, (uint32_t)(sw0 + ln), a.k0, a.k1);
        int f0, f1 = -1, b = 0;
        if ((uint64_t)r.x < a.thr && d.npb > 0) {
            b = (int)bounded(r.y, (uint32_t)d.npb);
            f0 = l.pb[2 * b];
            f1 = l.pb[2 * b + 1];
        } else {
            f0 = (int)bounded(r.y, (uint32_t)d.N);
        }
        pcode = f0 | ((f1 + 1) << 8) | (b << 16);
        plu = __log2f(uniform<float>(r.z, r.w));
    };

    uint32_t acc = 0, prop = 0;
    int64_t k = 0;
    for (int64_t step = 0; step < total; step++) {
        // slots sw0 .. sw0 + 63 are drawn before their proposals run; a pair
        // takes two slots, so a chunk's last pair may step past its end
        for (int sw = 0; sw < d.N;) {
          const int sw0 = sw & ~63, swe = min(sw0 + 64, d.N);
          draw(step, sw0);
          for (; sw < swe; sw++) {
            const int code = __builtin_amdgcn_readlane(pcode, sw - sw0);
            const float lu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(plu), sw - sw0));
            const int f0 = code & 0xFF, f1 = ((code >> 8) & 0xFF) - 1, b = code >> 16;
            prop++;
            // the proposal's spins and bias term are wave-uniform: integer
            // selects and sign flips keep them in scalar registers
            const int bit0 = wbit(st0, st1, f0) ? 1 : 0;
            const int bit1 = f1 >= 0 && wbit(st0, st1, f1) ? 1 : 0;
            const float s0 = sgn_of(bit0);
            float bre = flip_if(rfl(l.rb[f0].x), bit0);  // s0 Re b_f0
            float lr;
            const int combo = f1 >= 0 ? bit0 ^ bit1 : 0;
            auto e1 = [&](int u) __attribute__((always_inline)) { return U.entry_a(d, l, u, f0); };
            auto ep = [&](int u) __attribute__((always_inline)) { return U.pentry_a(d, l, u, b, combo); };
            if (f1 < 0) {
                lr = U.lratio1(e1, -s0);
            } else {
                sw++;
                const float s1 = sgn_of(bit1);
                bre += flip_if(rfl(l.rb[f1].x), bit1);
                if constexpr (PAIRS) {
                    lr = U.lratio1(ep, -s0);
                } else {
                    lr = U.lratio2(d, l, f0, f1, -s0, -s1);
                }
            }
            // log2 |psi'/psi|^2 = wave sum of the lanes' logs + 2 Re(bias
            // change) / ln 2, the bias change -2 sum s_f Re b_f = -2 bre
            const float lr2 = wave_sum(lr) - 5.7707801635558536f * bre;
            if (lu < lr2) {
                if (PAIRS || f1 < 0) {
                    U.accept1(f1 < 0 ? wofs1 : wofs2, s0);
                } else {
                    U.accept2(d, l, f0, f1, s0, sgn_of(bit1));
                }
                if (f0 < 64) st0 ^= (uint64_t)1 << f0;
                else st1 ^= (uint64_t)1 << (f0 - 64);
                if (f1 >= 0) {
                    if (f1 < 64) st0 ^= (uint64_t)1 << f1;
                    else st1 ^= (uint64_t)1 << (f1 - 64);
                }
                acc++;
            }
          }
        }
        U.resync();
        if (step >= a.warmup && (step - a.warmup) % a.step == 0) {
            const int64_t smp = off + k;
            if (ln == 0) {
                a.states[smp * d.nw] = st0;
                if (d.nw > 1) a.states[smp * d.nw + 1] = st1;
            }
#pragma unroll
            for (int u = 0; u < K; u++)
                if (u < K - 1 || U.vlast) a.Tout[smp * d.H + ln + u * 64] = U.tn[u];
            if (!a.obs_later) wave_obs<K, SYM>(d, l, U, st0, st1, a.npair, a.obs + smp * 2 * a.T.n_groups);
            k++;
        }
    }
    if (ln == 0) {
        a.acc[chain] = acc;
        a.prop[chain] = prop;
    }
}

}  // namespace rbmgpu
