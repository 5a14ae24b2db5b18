// kern_learn.hpp -- Metropolis-Hastings with a learned cluster proposal and
// the training of its network.  No counterpart in the reference, whose
// proposals are single and bond flips (metropolis_sampler.cpp:195-219).
//
// Proposal network: logits z(s) = W2 relu(W1 s + b1) + b2 over the N sites
// (s_j = +-1).  A learned move draws a seed site i ~ softmax(z(sigma)) and a
// size k ~ U{1..kmax}, and flips the cluster C(i, k) = the first k sites of
// the lattice loop through i (hexagons / vertex stars).  Its proposal
// probability is q(F | sigma) = (1 / kmax) sum_{j in seeds(F)} p_j(sigma)
// over the seeds generating the same set, so the Hastings factor is
// sum_j p_j(sigma') / sum_j p_j(sigma).  With probability 1 - lambda a slot
// makes the plain move instead (single site or bond, as the reference);
// the mixture of the two kernels is a valid Metropolis-Hastings chain.
//
// Lockstep blocks.  A block holds 8 chains, one per wavefront.  Whether a
// slot makes a learned or a plain move is drawn from a Philox stream shared
// by the block (it does not depend on any chain's state, so each chain's
// mixture of the two Metropolis-Hastings kernels stays valid); only the
// learned slots evaluate the network, for all 8 chains at once with two
// bf16 MFMA GEMMs (v_mfma_f32_16x16x16_bf16), H = S W1^T and Z = relu(H)
// W2^T, and only they synchronise the block.  Small networks (hidden width
// and sites at most 64) skip the batching: each wavefront evaluates its own
// chain's network on the FP32 VALU and the block never synchronises in the
// slot loop.  A chain's current logits go stale when it accepts a plain move
// and are recomputed at the next learned slot.  Logits are always computed from the state itself (no incremental
// hidden layer), so q is an exact function of the state and the chain is
// exactly balanced for the network as evaluated.
//
// bf16 MFMA operand layout (16x16x16): lane l supplies A[l%16][4(l/16)+j]
// and B[4(l/16)+j][l%16], j < 4; accumulator r of lane l is D[4(l/16)+r][l%16].
#pragma once
#include "kern_impl.hpp"

namespace rbmgpu {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f_l __attribute__((ext_vector_type(4)));

__device__ inline uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

struct MlpLds {
    uint16_t *W1, *W2, *Sg, *Ha;
    float *b1, *b2, *Zt, *Zc;
    uint8_t *cyc, *seedc, *seedl;
    float *W1t, *W2t, *Hw;  // per-wavefront forward: W1^T [site][h], W2^T [h][site] (FP32), hidden rows
};

// the network region after the sampler tables (layout of mlp_lds_bytes)
__device__ inline MlpLds mlp_stage(const LearnArgs& a, char* base) {
    const int Hn = a.Hn, Np = a.Np, N = a.s.d.N, kmax = a.kmax;
    MlpLds m;
    char* at = base;
    m.W1 = reinterpret_cast<uint16_t*>(at);
    at += (size_t)Hn * Np * 2;
    m.W2 = reinterpret_cast<uint16_t*>(at);
    at += (size_t)Hn * Np * 2;
    m.b1 = reinterpret_cast<float*>(at);
    at += (size_t)Hn * 4;
    m.b2 = reinterpret_cast<float*>(at);
    at += (size_t)Np * 4;
    m.Sg = reinterpret_cast<uint16_t*>(at);
    at += 16 * (size_t)Np * 2;
    m.Ha = reinterpret_cast<uint16_t*>(at);
    at += 16 * (size_t)Hn * 2;
    m.Zt = reinterpret_cast<float*>(at);
    at += 16 * (size_t)Np * 4;
    m.Zc = reinterpret_cast<float*>(at);
    at += 16 * (size_t)Np * 4;
    m.cyc = reinterpret_cast<uint8_t*>(at);
    at += (size_t)N * kmax;
    m.seedc = reinterpret_cast<uint8_t*>(at);
    at += (size_t)N * kmax;
    m.seedl = reinterpret_cast<uint8_t*>(at);
    at += (size_t)6 * N * kmax;
    m.W1t = m.W2t = m.Hw = nullptr;
    if (mlp_wave_fwd(Hn, N)) {
        at = base + (((size_t)(at - base) + 15) & ~(size_t)15);
        m.W1t = reinterpret_cast<float*>(at);
        m.W2t = m.W1t + (size_t)Hn * Np;
        m.Hw = m.W2t + (size_t)Hn * Np;
        for (int i = threadIdx.x; i < Hn * Np; i += blockDim.x) {
            const int h = i / Np, j = i - h * Np;  // W1 row h, W2 row j = i / Hn
            m.W1t[j * Hn + h] = __uint_as_float((uint32_t)a.W1[i] << 16);
            const int j2 = i / Hn, h2 = i - j2 * Hn;
            m.W2t[h2 * Np + j2] = __uint_as_float((uint32_t)a.W2[i] << 16);
        }
    }
    for (int i = threadIdx.x; i < Hn * Np / 2; i += blockDim.x) {
        reinterpret_cast<uint32_t*>(m.W1)[i] = reinterpret_cast<const uint32_t*>(a.W1)[i];
        reinterpret_cast<uint32_t*>(m.W2)[i] = reinterpret_cast<const uint32_t*>(a.W2)[i];
    }
    for (int i = threadIdx.x; i < Hn; i += blockDim.x) m.b1[i] = a.b1[i];
    for (int i = threadIdx.x; i < Np; i += blockDim.x) m.b2[i] = a.b2[i];
    for (int i = threadIdx.x; i < N * kmax; i += blockDim.x) {
        m.cyc[i] = a.cyc[i];
        m.seedc[i] = a.seedc[i];
    }
    for (int i = threadIdx.x; i < 6 * N * kmax; i += blockDim.x) m.seedl[i] = a.seedl[i];
    for (int i = threadIdx.x; i < 16 * Np; i += blockDim.x) m.Sg[i] = 0;
    __syncthreads();
    return m;
}

// Zt = W2 relu(W1 Sg + b1) + b2 for the 16 rows of Sg (all waves of the
// block take part, 16-column tiles dealt round-robin; syncs included)
__device__ inline void mlp_forward16(const MlpLds& m, int Hn, int Np) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    const int nwv = (int)(blockDim.x >> 6);
    __syncthreads();  // Sg rows written
    for (int ct = w; ct < Hn / 16; ct += nwv) {
        v4f_l acc = {0, 0, 0, 0};
        for (int kb = 0; kb < Np / 16; kb++) {
            v4s av = *reinterpret_cast<const v4s*>(m.Sg + lr * Np + 16 * kb + 4 * lq);
            v4s bv = *reinterpret_cast<const v4s*>(m.W1 + (16 * ct + lr) * Np + 16 * kb + 4 * lq);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
        const int h = 16 * ct + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) m.Ha[(4 * lq + r) * Hn + h] = to_bf16(fmaxf(acc[r] + m.b1[h], 0.f));
    }
    __syncthreads();
    for (int ct = w; ct < Np / 16; ct += nwv) {
        v4f_l acc = {0, 0, 0, 0};
        for (int kb = 0; kb < Hn / 16; kb++) {
            v4s av = *reinterpret_cast<const v4s*>(m.Ha + lr * Hn + 16 * kb + 4 * lq);
            v4s bv = *reinterpret_cast<const v4s*>(m.W2 + (16 * ct + lr) * Hn + 16 * kb + 4 * lq);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
        const int o = 16 * ct + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) m.Zt[(4 * lq + r) * Np + o] = acc[r] + m.b2[o];
    }
    __syncthreads();
}

// this wavefront's logits Zt[w] = W2 relu(W1 s + b1) + b2 for its chain's
// state with the flips f0, f1 applied, on the FP32 VALU (the bf16 weights
// as FP32): lane h sums its hidden unit over the sites, the hidden row goes
// through LDS, lane j sums its logit over the hidden units; wavefront-local
// (no block synchronisation)
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ inline void mlp_forward_wave(const MlpLds& m, int Hn, int Np, int N, const uint64_t* st, uint64_t f0,
                                        uint64_t f1, int w, int lane) {
    float* hw = m.Hw + w * Hn;
    if (lane < Hn) {
        float acc = m.b1[lane];
#pragma unroll 8
        for (int i = 0; i < N; i++) {
            const int b = bitof(st, i) ^ (int)(((i < 64 ? f0 : f1) >> (i & 63)) & 1);
            const float wv = m.W1t[i * Hn + lane];
            acc = b ? acc - wv : acc + wv;
        }
        hw[lane] = fmaxf(acc, 0.f);
    }
    wave_lds_sync();
    if (lane < N) {
        float acc = m.b2[lane];
#pragma unroll 8
        for (int h = 0; h < Hn; h++) acc = fmaf(m.W2t[h * Np + lane], hw[h], acc);
        m.Zt[w * Np + lane] = acc;
    }
    wave_lds_sync();
}

// this wave's row of Sg from a state with the flips of mask applied
__device__ inline void mlp_put_state(const MlpLds& m, int Np, int N, const uint64_t* st, uint64_t f0, uint64_t f1,
                                     int w, int lane) {
    for (int j = lane; j < Np; j += 64) {
        uint16_t v = 0;
        if (j < N) {
            int b = bitof(st, j) ^ (int)(((j < 64 ? f0 : f1) >> (j & 63)) & 1);
            v = b ? 0xBF80 : 0x3F80;  // bf16 -1, +1
        }
        m.Sg[w * Np + j] = v;
    }
}

// log-sum-exp of a row of logits (sites j < N), uniform over the wave
__device__ inline float row_lse(const float* z, int N, int lane) {
    float z0 = lane < N ? z[lane] : -INFINITY, z1 = lane + 64 < N ? z[lane + 64] : -INFINITY;
    float mx = group_max<64>(fmaxf(z0, z1));
    float e = (lane < N ? __expf(z0 - mx) : 0.f) + (lane + 64 < N ? __expf(z1 - mx) : 0.f);
    return mx + __logf(group_sum<64>(e));
}

// sum of softmax probabilities of the seeds of cluster (i, k)
__device__ inline float seed_mass(const MlpLds& m, const float* z, float lse, int i, int k, int kmax) {
    const int c = m.seedc[i * kmax + k - 1];
    float q = 0;
    for (int t = 0; t < c; t++) q += __expf(z[m.seedl[(i * kmax + k - 1) * 6 + t]] - lse);
    return q;
}

// WPE: minimum waves per SIMD the register allocation must allow (1: no
// limit, the fastest code for one block per CU, blocks of at most
// kMlpChains chains; 4: at most 128 VGPRs, so that two blocks of 8 or one
// of 16 chains share a CU)
template <int K, int WPE>
__global__ void __launch_bounds__(64 * (WPE >= 4 ? kMlpChainsMax : kMlpChains))
    __attribute__((amdgpu_waves_per_eu(WPE))) k_sample_mlp(LearnArgs la) {
    extern __shared__ __align__(16) char smem[];
    const SampArgs<float>& a = la.s;
    const Dims d = a.d;
    Lds<float> l = stage<float>(d, a.t, a.T, a.pb, smem);
    MlpLds m = mlp_stage(la, smem + lds_bytes<float>(d));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int chain = blockIdx.x * (int)(blockDim.x >> 6) + w;  // blocks of 1..kMlpChainsMax chains
    const bool live = chain < a.n_chains;
    const int N = d.N, Np = la.Np, Hn = la.Hn, kmax = la.kmax;
    int64_t spc = a.n_samples / a.n_chains, res = a.n_samples - spc * a.n_chains;
    int64_t nc = live ? spc + (chain == 0 ? res : 0) : 0;
    int64_t off = chain == 0 ? 0 : chain * spc + res;
    // every chain of the block runs the block's longest schedule (lockstep)
    int64_t total = (spc + (blockIdx.x == 0 ? res : 0)) * a.step + a.warmup;

    const uint32_t epoch = a.epoch_dev ? *a.epoch_dev : a.epoch;
    uint64_t st[2] = {0, 0};
    {
        u4 r = philox((uint32_t)chain, epoch, 0xFFFFFFFFu, 0u, a.k0, a.k1);
        uint32_t wd[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int q = 0; q < 4; q++) st[q >> 1] |= (uint64_t)wd[q] << (32 * (q & 1));
        if (N < 64) st[0] &= (((uint64_t)1) << N) - 1;
        if (N <= 64) st[1] = 0;
        else if (N < 128) st[1] &= (((uint64_t)1) << (N - 64)) - 1;
    }
    Units<float, 64, K> U;
    U.init(d, lane);
    U.from_state(d, l, a.t.hb, st);
    cplx<float> pv[16];
    const bool wf = mlp_wave_fwd(Hn, N);  // per-wavefront forward passes (uniform)
    // logits of the initial states
    if (wf) {
        mlp_forward_wave(m, Hn, Np, N, st, 0, 0, w, lane);
    } else {
        mlp_put_state(m, Np, N, st, 0, 0, w, lane);
        mlp_forward16(m, Hn, Np);
    }
    for (int j = lane; j < Np; j += 64) m.Zc[w * Np + j] = m.Zt[w * Np + j];
    if (wf) wave_lds_sync();
    float lse = row_lse(m.Zt + w * Np, N, lane);

    // per-slot randomness, drawn ahead by the lanes (lane j: slot sw0 + j):
    // A = philox(chain, epoch, step, slot), B = philox(.., slot | 2^31), and
    // the block's learned / plain choice G = philox(2^30 | block, epoch,
    // step, slot)
    int pcode = 0;
    float pua = 0, pus = 0;
    auto draw = [&](int64_t step, int sw0) {
        uint32_t slot = (uint32_t)(sw0 + lane);
        u4 A = philox((uint32_t)chain, epoch, (uint32_t)step, slot, a.k0, a.k1);
        u4 B = philox((uint32_t)chain, epoch, (uint32_t)step, slot | 0x80000000u, a.k0, a.k1);
        u4 G = philox(0x40000000u | (uint32_t)blockIdx.x, epoch, (uint32_t)step, slot, a.k0, a.k1);
        if (G.x < la.lam_thr) {
            pcode = (int)(0x40000000u | ((1u + bounded(B.y, (uint32_t)kmax)) << 16));
        } else {
            int f0, f1 = -1;
            if ((uint64_t)A.y < a.thr && d.npb > 0) {
                int bi = (int)bounded(A.z, (uint32_t)d.npb);
                f0 = l.pb[2 * bi];
                f1 = l.pb[2 * bi + 1];
            } else {
                f0 = (int)bounded(A.z, (uint32_t)N);
            }
            pcode = f0 | ((f1 + 1) << 8);
        }
        pua = uniform<float>(A.w, 0);
        pus = uniform<float>(B.x, 0);
    };

    uint32_t acc = 0, prop = 0, lacc = 0, lprop = 0;
    bool stale = false;  // Zc, lse are not those of the current state
    uint64_t rs0 = 0, rs1 = 0;
    int rseed = -1, rk = 0;
    float ra = 0;
    int64_t k = 0;
    for (int64_t step = 0; step < total; step++) {
        int sw0 = -64;
        for (int sw = 0; sw < N; sw++) {
            if (sw >= sw0 + 64) {
                sw0 = sw;
                draw(step, sw0);
            }
            const int code = __builtin_amdgcn_readlane(pcode, sw - sw0);
            const float ua = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pua), sw - sw0));
            const float us = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pus), sw - sw0));
            const bool learned = (code & 0x40000000) != 0;
            int fl[8];
            float sg[8];
            int mask = 0, seed = 0, kk = 0;
            uint64_t fm0 = 0, fm1 = 0;
            if (learned) {
                // the block's chains whose current logits are stale (an
                // accepted plain move since their last evaluation)
                if (wf) {
                    if (stale) {
                        mlp_forward_wave(m, Hn, Np, N, st, 0, 0, w, lane);
                        for (int j = lane; j < Np; j += 64) m.Zc[w * Np + j] = m.Zt[w * Np + j];
                        wave_lds_sync();
                        lse = row_lse(m.Zt + w * Np, N, lane);
                        stale = false;
                    }
                } else if (__syncthreads_or(stale ? 1 : 0)) {
                    mlp_put_state(m, Np, N, st, 0, 0, w, lane);
                    mlp_forward16(m, Hn, Np);
                    if (stale) {
                        for (int j = lane; j < Np; j += 64) m.Zc[w * Np + j] = m.Zt[w * Np + j];
                        lse = row_lse(m.Zt + w * Np, N, lane);
                        stale = false;
                    }
                }
                // seed i ~ softmax(z(sigma)): inclusive scans of the lanes'
                // probabilities (sites lane and lane + 64), first crossing of us
                const float* zc = m.Zc + w * Np;
                // (sites 64.. only when the first 64 do not cross us)
                float c0 = lane < N ? __expf(zc[lane] - lse) : 0.f;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    float t0 = __shfl_up(c0, o);
                    if (lane >= o) c0 += t0;
                }
                const uint64_t b0 = __ballot(lane < N && c0 > us);
                if (b0) {
                    seed = __ffsll((long long)b0) - 1;
                } else if (N > 64) {
                    float c1 = lane + 64 < N ? __expf(zc[lane + 64] - lse) : 0.f;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        float t1 = __shfl_up(c1, o);
                        if (lane >= o) c1 += t1;
                    }
                    const float tot0 = __shfl(c0, 63);
                    const uint64_t b1 = __ballot(lane + 64 < N && tot0 + c1 > us);
                    seed = b1 ? 64 + __ffsll((long long)b1) - 1 : N - 1;
                } else {
                    seed = N - 1;
                }
                kk = (code >> 16) & 0xFF;
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    fl[t] = t < kk ? (int)m.cyc[seed * kmax + t] : 0;
                    if (t < kk) mask |= 1 << t;
                }
            } else {
                int f0 = code & 0xFF, f1 = (code >> 8) - 1;
                fl[0] = f0;
                fl[1] = f1 >= 0 ? f1 : 0;
                mask = f1 >= 0 ? 3 : 1;
#pragma unroll
                for (int t = 2; t < 8; t++) fl[t] = 0;
            }
#pragma unroll
            for (int t = 0; t < 8; t++) {
                sg[t] = bitof(st, fl[t]) ? -1.f : 1.f;
                if ((mask >> t) & 1) {
                    if (fl[t] < 64) fm0 |= (uint64_t)1 << fl[t];
                    else fm1 |= (uint64_t)1 << (fl[t] - 64);
                }
            }
            prop++;
            // log2 |psi'/psi|^2 of the RBM
            cplx<float> p;
            int pe;
            U.ratio_mask(d, l, fl, sg, mask, p, pe);
            float lr2 = group_sum<64>(M<float>::log2(norm2(p)) + (float)(2 * pe));
            float bre = 0;
#pragma unroll
            for (int t = 0; t < 8; t++)
                if ((mask >> t) & 1) bre += -2.f * sg[t] * l.rb[fl[t]].re;
            lr2 += 2.8853900817779268f * bre;  // 2 / ln 2
            // logits of the proposed state (learned slots, lockstep)
            const float* zt = m.Zt + w * Np;
            float lse2 = 0.f;
            if (learned) {
                if (wf) {
                    mlp_forward_wave(m, Hn, Np, N, st, fm0, fm1, w, lane);
                } else {
                    mlp_put_state(m, Np, N, st, fm0, fm1, w, lane);
                    mlp_forward16(m, Hn, Np);
                }
                lse2 = row_lse(zt, N, lane);
                float qf = seed_mass(m, m.Zc + w * Np, lse, seed, kk, kmax);
                float qr = seed_mass(m, zt, lse2, seed, kk, kmax);
                lr2 += M<float>::log2(qr) - M<float>::log2(qf);
                lprop++;
                rs0 = st[0];
                rs1 = st[1];
                rseed = seed;
                rk = kk;
                ra = fminf(1.f, M<float>::exp2(lr2));
            }
            if (ua < M<float>::exp2(lr2)) {
                U.accept_mask(d, l, fl, sg, mask);
                st[0] ^= fm0;
                st[1] ^= fm1;
                if (learned) {
                    for (int j = lane; j < Np; j += 64) m.Zc[w * Np + j] = zt[j];
                    if (wf) wave_lds_sync();
                    lse = lse2;
                    lacc++;
                } else {
                    stale = true;
                }
                acc++;
            }
        }
        if (step >= a.warmup && (step - a.warmup) % a.step == 0) {
            if (k < nc) {
                int64_t smp = off + k;
                if (lane == 0) {
                    for (int q = 0; q < d.nw; q++) a.states[smp * d.nw + q] = st[q];
                    la.rec_st[smp * d.nw] = rs0;
                    if (d.nw > 1) la.rec_st[smp * d.nw + 1] = rs1;
                    la.rec_seed[smp] = (int16_t)rseed;
                    la.rec_k[smp] = (uint8_t)rk;
                    la.rec_a[smp] = ra;
                }
                rseed = -1;
#pragma unroll
                for (int u = 0; u < K; u++)
                    if (U.a[u] >= 0) a.Tout[smp * d.H + lane + u * 64] = U.tn[u];
                eval_obs<float, 64, K, false>(d, l, U, pv, st, lane, a.obs + smp * 2 * a.T.n_groups);
            }
            k++;
        }
    }
    if (live && lane == 0) {
        a.acc[chain] = acc;
        a.prop[chain] = prop;
        la.lacc[chain] = lacc;
        la.lprop[chain] = lprop;
    }
}

// logits of host-given states through the sampler's own forward (the
// per-wavefront FP32 one for small networks, else the bf16 MFMA one; 8 per
// block)
__global__ void __launch_bounds__(64 * kMlpChains) k_mlp_logits(LearnArgs la, const uint64_t* states, int64_t n, float* z) {
    extern __shared__ __align__(16) char smem[];
    MlpLds m = mlp_stage(la, smem);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, N = la.s.d.N, nw = la.s.d.nw;
    const int64_t r = (int64_t)blockIdx.x * kMlpChains + w;
    uint64_t st[2] = {0, 0};
    if (r < n)
        for (int q = 0; q < nw; q++) st[q] = states[r * nw + q];
    if (mlp_wave_fwd(la.Hn, N)) {
        mlp_forward_wave(m, la.Hn, la.Np, N, st, 0, 0, w, lane);
    } else {
        mlp_put_state(m, la.Np, N, st, 0, 0, w, lane);
        mlp_forward16(m, la.Hn, la.Np);
    }
    if (r < n)
        for (int j = lane; j < N; j += 64) z[r * N + j] = m.Zt[w * la.Np + j];
}

// ------------------------------------------------------------- training --
// Per record r (one block of Hn threads): s, u = W1 s + b1, a = relu(u),
// z = W2 a + b2, p = softmax(z), dL/dz = inv_b [w_r (p - e_seed) +
// beta p (ln p + H)], dL/du = (W2^T dz) [u > 0], and the record's loss
// -w_r ln p_seed - beta H.  Records with seed < 0 contribute nothing.
__global__ void __launch_bounds__(512) k_train_fwd(TrainArgs a) {
    __shared__ float sv[128], av[512], zv[128], red[512];
    const int64_t r = blockIdx.x;
    const int t = threadIdx.x, N = a.N, Hn = a.Hn;
    const int seed = a.seed[r];
    for (int j = t; j < N; j += blockDim.x) {
        float v = ((a.st[r * a.nw + (j >> 6)] >> (j & 63)) & 1) ? -1.f : 1.f;
        sv[j] = v;
        a.S[r * N + j] = seed >= 0 ? v : 0.f;
    }
    __syncthreads();
    float u = 0;
    if (t < Hn) {
        u = a.b1[t];
        for (int j = 0; j < N; j++) u += a.W1[t * N + j] * sv[j];
        av[t] = fmaxf(u, 0.f);
        a.A[r * Hn + t] = seed >= 0 ? av[t] : 0.f;
    }
    __syncthreads();
    float zo = -INFINITY;
    if (t < N) {
        zo = a.b2[t];
        for (int h = 0; h < Hn; h++) zo += a.W2[t * Hn + h] * av[h];
    }
    // softmax over the N logits
    red[t] = zo;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = fmaxf(red[t], red[t + s]);
        __syncthreads();
    }
    const float mx = red[0];
    __syncthreads();
    red[t] = t < N ? __expf(zo - mx) : 0.f;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const float lse = mx + __logf(red[0]);
    __syncthreads();
    float lp = zo - lse, pp = t < N ? __expf(lp) : 0.f;
    red[t] = t < N ? -pp * lp : 0.f;  // entropy terms
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const float H = red[0];
    const float wr = seed >= 0 ? a.w[r] : 0.f;
    if (t < N) {
        float dz = seed >= 0 ? a.inv_b * (wr * (pp - (t == seed ? 1.f : 0.f)) + a.beta * pp * (lp + H)) : 0.f;
        zv[t] = dz;
        a.dZ[r * N + t] = dz;
        if (t == seed) a.loss[r] = -wr * lp - a.beta * H;
    }
    if (t == 0 && seed < 0) a.loss[r] = 0.f;
    __syncthreads();
    if (t < Hn) {
        float da = 0;
        for (int o = 0; o < N; o++) da += a.W2[o * Hn + t] * zv[o];
        a.dU[r * Hn + t] = u > 0 ? da : 0.f;
    }
}

// One batch slice z (records z kc .. z kc + kc - 1) of C = X^T Y (M x Nn;
// X: K x M, Y: K x Nn, row-major) into part + z ldz, 64 x 64 tiles through
// LDS; the blocks of the first column tile also sum X's columns of the
// slice into part + z ldz + soff (the bias gradients)
__global__ void __launch_bounds__(256) k_gemm_tn_part(const float* X, int ldx, const float* Y, int ldy, float* part,
                                                      int64_t ldz, int64_t soff, int M, int Nn, int64_t K,
                                                      int64_t kc) {
    __shared__ float xs[16][64], ys[16][64];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int64_t kb = (int64_t)blockIdx.z * kc, ke = kb + kc < K ? kb + kc : K;
    float c[4][4] = {};
    float cs = 0.f;
    for (int64_t k0 = kb; k0 < ke; k0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            int kk = e / 64, j = e % 64;
            int64_t kr = k0 + kk;
            xs[kk][j] = (kr < ke && m0 + j < M) ? X[kr * ldx + m0 + j] : 0.f;
            ys[kk][j] = (kr < ke && n0 + j < Nn) ? Y[kr * ldy + n0 + j] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) c[i][j] += xs[kk][ty * 4 + i] * ys[kk][tx * 4 + j];
        }
        if (blockIdx.x == 0 && threadIdx.x < 64)
#pragma unroll
            for (int kk = 0; kk < 16; kk++) cs += xs[kk][threadIdx.x];
        __syncthreads();
    }
    float* C = part + (int64_t)blockIdx.z * ldz;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int mm = m0 + ty * 4 + i, nn = n0 + tx * 4 + j;
            if (mm < M && nn < Nn) C[(int64_t)mm * Nn + nn] = c[i][j];
        }
    if (blockIdx.x == 0 && threadIdx.x < 64 && m0 + (int)threadIdx.x < M) C[soff + m0 + threadIdx.x] = cs;
}

// out[i] = sum over the ks slices of part[z ldz + i], in slice order
__global__ void __launch_bounds__(256) k_sum_slices(const float* part, int ks, int64_t ldz, int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < ks; z++) s += part[(int64_t)z * ldz + i];
    out[i] = s;
}

}  // namespace rbmgpu
