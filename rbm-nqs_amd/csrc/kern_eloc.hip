// kern_eloc.hip -- local observables of stored samples, one thread per
// (sample, term): E_loc(s) = sum_t sum_(s' in row t of s) <s|h_t|s'>
// psi(s')/psi(s) (local_op.cpp:31-65, rbm_base.cpp:145-200), FP32 basic
// RBM, terms of at most four sites.
//
// The sampler stores every kept state and its tanh(theta) row (S x H), so
// the ratios need no state of the chain: for a term flipping the sites F,
// psi(s')/psi(s) = exp(-2 sum_f sigma_f c_f) prod_j (ch_j + tanh(theta_j)
// sh_j), where (ch_j, sh_j) = (cosh, sinh) of -2 sum_f sigma_f W_fj come
// from the cosh(2W) / sinh(2W) tables by the addition theorem, one site at
// a time in the order of the term's sites -- the same recurrence as the
// sampler's lane-parallel evaluation, taken here over all H units by one
// thread instead of over the lanes of a chain.  A thread holds no
// cross-lane state, so the launch has every sample's terms in flight at
// once instead of the sampler's two wavefronts per SIMD.
#include <hip/hip_runtime.h>

#include "kern_api.hpp"

namespace rbmgpu {

namespace {

__device__ inline int state_bit(const uint64_t* st, int i) { return (int)((st[i >> 6] >> (i & 63)) & 1); }

// blockDim.x = kElocThreads; LDS: cosh(2W), sinh(2W) transposed to [j][f]
// (the lanes of a wavefront read one unit j of different sites: distinct
// banks), then one complex partial per thread
__global__ void __launch_bounds__(kElocThreads) k_eloc(ElocArgs a) {
    extern __shared__ __align__(16) char smem[];
    const int N = a.N, H = a.H, nt = a.T.n_terms;
    cplx<float>* Ct = reinterpret_cast<cplx<float>*>(smem);
    cplx<float>* St = Ct + (size_t)H * N;
    double* part = reinterpret_cast<double*>(St + (size_t)H * N);
    for (int i = threadIdx.x; i < H * N; i += blockDim.x) {
        const int j = i / N, f = i - j * N;
        Ct[i] = a.C[f * H + j];
        St[i] = a.S[f * H + j];
    }
    __syncthreads();
    const int nr = a.n_ratio;             // terms with an off-diagonal entry: one thread each
    const int per = kElocThreads / nr;  // samples per pass of the block
    const int q = threadIdx.x / nr, t = a.ratio_term[threadIdx.x - q * nr];
    for (int64_t s0 = (int64_t)blockIdx.x * per; s0 < a.n; s0 += (int64_t)gridDim.x * per) {
        const int64_t smp = s0 + q;
        const bool live = q < per && smp < a.n;
        double cr = 0, ci = 0;
        if (live) {
            uint64_t st[2] = {a.states[smp * a.nw], a.nw > 1 ? a.states[smp * a.nw + 1] : 0};
            const int k = a.T.k[t];
            int fs[4];
            float sg[4];
            int loc = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                fs[i] = i < k ? (int)a.T.sites[t * 8 + i] : 0;
                const int b = i < k ? state_bit(st, fs[i]) : 0;
                sg[i] = b ? -1.f : 1.f;
                loc |= b << i;
            }
            const int base = a.T.rowbase[t] + loc;
            const int e0 = a.T.rowptr[base], e1 = a.T.rowptr[base + 1];
            const cplx<float>* th = a.tanh + smp * H;
            for (int e = e0; e < e1; e++) {
                const int mask = a.T.mask[e];
                const double vr = a.T.val[2 * e], vi = a.T.val[2 * e + 1];
                if (mask == 0) {
                    cr += vr;
                    ci += vi;
                    continue;
                }
                bool on[4];
                cplx<float> bias{0.f, 0.f};
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    on[i] = (mask >> i) & 1;
                    if (on[i]) bias = bias + (-2 * sg[i]) * a.rb[fs[i]];
                }
                // sites not flipped enter as cosh = 1, sinh = 0 (exact)
                cplx<float> p{1.f, 0.f};
                int pe = 0;
                for (int j = 0; j < H; j++) {
                    cplx<float> ch{1.f, 0.f}, sh{0.f, 0.f};
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const cplx<float> c = on[i] ? Ct[j * N + fs[i]] : cplx<float>{1.f, 0.f};
                        const cplx<float> sn = on[i] ? St[j * N + fs[i]] : cplx<float>{0.f, 0.f};
                        const cplx<float> snu = (-sg[i]) * sn;
                        const cplx<float> nch = ch * c + sh * snu;
                        sh = sh * c + ch * snu;
                        ch = nch;
                    }
                    p = p * (ch + th[j] * sh);
                    if ((j & 3) == 3) renorm(p, pe);
                }
                renorm(p, pe);
                float sn, cs;
                sincos_f(bias.im, sn, cs);
                const float ex = __expf(bias.re);
                const cplx<float> r = cplx<float>{ex * cs, ex * sn} * p;
                const double sc = ::ldexp(1.0, pe), rr = (double)r.re * sc, ri = (double)r.im * sc;
                cr += vr * rr - vi * ri;
                ci += vr * ri + vi * rr;
            }
        }
        part[2 * threadIdx.x] = cr;
        part[2 * threadIdx.x + 1] = ci;
        __syncthreads();
        if (live && threadIdx.x - q * nr == 0) {
            // the sample's group sums, terms in order: a ratio term's sum from
            // its thread, a diagonal term's entries here
            uint64_t st[2] = {a.states[smp * a.nw], a.nw > 1 ? a.states[smp * a.nw + 1] : 0};
            double* out = a.obs + smp * 2 * a.T.n_groups;
            int g = a.T.group[0];
            double gr = 0, gi = 0;
            for (int ti = 0; ti < nt; ti++) {
                const int gt = a.T.group[ti];
                if (gt != g) {
                    out[2 * g] = gr;
                    out[2 * g + 1] = gi;
                    gr = gi = 0;
                    g = gt;
                }
                const int slot = a.term_slot[ti];
                if (slot >= 0) {
                    gr += part[2 * (q * nr + slot)];
                    gi += part[2 * (q * nr + slot) + 1];
                } else {
                    const int k = a.T.k[ti];
                    int loc = 0;
                    for (int i = 0; i < k; i++) loc |= state_bit(st, a.T.sites[ti * 8 + i]) << i;
                    const int base = a.T.rowbase[ti] + loc;
                    for (int e = a.T.rowptr[base]; e < a.T.rowptr[base + 1]; e++) {
                        gr += a.T.val[2 * e];
                        gi += a.T.val[2 * e + 1];
                    }
                }
            }
            out[2 * g] = gr;
            out[2 * g + 1] = gi;
        }
        __syncthreads();
    }
}

}  // namespace

size_t eloc_lds_bytes(int N, int H) { return (size_t)N * H * 16 + (size_t)kElocThreads * 16; }

// resident k_eloc blocks per CU (the grid is sized to one full wave of
// them: its grid-stride loop then gives every block the same share)
int eloc_blocks_per_cu(int N, int H) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_eloc, kElocThreads, eloc_lds_bytes(N, H)) != hipSuccess) {
        (void)hipGetLastError();  // clear it: the launch that follows reports its own status
        return 1;
    }
    return nb > 0 ? nb : 1;
}

hipError_t launch_eloc(const ElocArgs& a, int blocks, hipStream_t st) {
    const size_t lds = eloc_lds_bytes(a.N, a.H);
    if (a.T.n_terms < 1 || a.n_ratio < 1 || a.n_ratio > kElocThreads || lds > 65536 || a.nw < 1 || a.nw > 2)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_eloc, dim3(blocks), dim3(kElocThreads), lds, st, a);
    return hipGetLastError();
}

}  // namespace rbmgpu
