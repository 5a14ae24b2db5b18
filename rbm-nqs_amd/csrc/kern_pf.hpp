// kern_pf.hpp -- the Pfaffian factor of the wave function (pfaffian.cpp)
// for one Markov chain per wavefront, N <= 32 spins.
//
// Register layout.  Lane l holds row i = l & 31 of the inverse of the skew
// matrix A(sigma), columns 16h .. 16h + 15 with h = l >> 5: v[q] =
// inv(i, 16h + q).  Rows i >= N are zero.
//
// Per-wave LDS scratch: X (N x N complex double, column-major X[j*N + i])
// for fresh inverses / Pfaffians, which are always formed in FP64, and as
// exchange space, D (8 x 32) for the Delta-a rows of a flip set, G (16 x 16)
// for the small 2m x 2m matrix.
//
// Ratio and update follow pfaffian.cpp:299-397 (Sherman-Morrison-Woodbury):
// for flips f_0..f_{m-1}, B = [Delta-a columns | unit columns e_{f_k}],
// Cinv' = Cinv0 + (G - G^T) / 2 with G = B^T inv B, ratio = sign Pf(Cinv'),
// and on acceptance inv += t2 - t2^T with t2 = (inv B) C (inv B)^T / 2,
// C = antisym(Cinv'^-1).
#pragma once
#include "dev_common.hpp"

namespace rbmgpu {

constexpr int kPfMaxN = 32;
constexpr int kPfMaxM = 8;  // flips per ratio

// cross-lane LDS ordering within one wavefront
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// scratch per wave in complex R entries: X (room for N x N complex double,
// 256 entries, or 4 rows accumulators), D (8 x 32), G (16 x 16), and 16
// ints for the flip list (8 complex slots)
template <typename R>
__host__ __device__ inline int pf_xsize(int N, int rows) {
    int x = N * N * (int)(sizeof(cplx<double>) / sizeof(cplx<R>));
    if (x < 256) x = 256;
    return 4 * rows > x ? 4 * rows : x;
}
template <typename R>
__host__ __device__ inline int pf_scratch(int N, int rows) {
    return pf_xsize<R>(N, rows) + 8 * 32 + 16 * 16 + 8;
}

template <typename R>
struct PfLds {
    const int16_t* bs;   // N x N, bs[i*N + j] = index-map entry b(i, j)
    const cplx<R>* fs;   // rows x 4: fs[r*4 + c]
    int N, rows;
    cplx<R>* X;          // this wave's exchange / accumulator scratch
    cplx<double>* Xd;    // the same region as N x N complex double
    cplx<R>* D;          // 8 x 32
    cplx<R>* G;          // 16 x 16
    int* F;              // 16 ints: the flip list of the current ratio
};

__device__ inline int bit2(const uint64_t* st, int i) { return (int)((st[i >> 6] >> (i & 63)) & 1); }

// A_ij for spin bits si, sj (pfaffian.cpp:123-137)
template <typename R>
__device__ inline cplx<R> pf_a(const PfLds<R>& l, int i, int j, int si, int sj) {
    if (i == j) return {0, 0};
    int bb = l.bs[i * l.N + j];
    int sp = i > j ? (si << 1) | sj : (sj << 1) | si;
    cplx<R> v = l.fs[(abs(bb) - 1) * 4 + sp];
    return bb < 0 ? cplx<R>{-v.re, -v.im} : v;
}

template <typename R>
__device__ inline cplx<R> cdiv(cplx<R> a, cplx<R> b) {
    R d = 1 / (b.re * b.re + b.im * b.im);
    return {(a.re * b.re + a.im * b.im) * d, (a.im * b.re - a.re * b.im) * d};
}

template <typename R>
__device__ inline cplx<R> shfl_c(cplx<R> v, int src) {
    return {__shfl(v.re, src), __shfl(v.im, src)};
}

// wave argmax of key over lanes (ties: smaller lane); returns the lane
__device__ inline int wave_argmax(double key, int lane) {
    int best = lane;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        double k2 = __shfl_xor(key, o);
        int b2 = __shfl_xor(best, o);
        if (k2 > key || (k2 == key && b2 < best)) {
            key = k2;
            best = b2;
        }
    }
    return __builtin_amdgcn_readfirstlane(best);
}

// Xd = A(sigma) (column-major), all N x N entries, in FP64
template <typename R>
__device__ inline void pf_build(const PfLds<R>& l, const uint64_t* st, int lane) {
    const int N = l.N;
    for (int e = lane; e < N * N; e += 64) {
        int i = e % N, j = e / N;
        cplx<R> a = pf_a(l, i, j, bit2(st, i), bit2(st, j));
        l.Xd[e] = {(double)a.re, (double)a.im};
    }
    wave_sync();
}

// Gauss-Jordan inverse of Xd in place (FP64) with partial pivoting (row swaps,
// columns unscrambled at the end); returns false if singular
template <typename R>
__device__ inline bool pf_gj(const PfLds<R>& l, int lane, int& ipiv_lane) {
    const int N = l.N;
    cplx<double>* X = l.Xd;
    bool ok = true;
    for (int c = 0; c < N; c++) {
        double key = -1;
        if (lane < N && lane >= c) {
            cplx<double> x = X[c * N + lane];
            key = (double)x.re * x.re + (double)x.im * x.im;
        }
        int p = wave_argmax(key, lane);
        if (lane == c) ipiv_lane = p;
        if (p != c) {
            if (lane < N) {
                cplx<double> a = X[lane * N + p], b = X[lane * N + c];
                X[lane * N + p] = b;
                X[lane * N + c] = a;
            }
            wave_sync();
        }
        cplx<double> piv = X[c * N + c];
        if (piv.re == 0 && piv.im == 0) ok = false;
        cplx<double> inv = cdiv(cplx<double>{1, 0}, piv);
        // row c: X(c, c) := 1, then row c *= 1 / piv
        if (lane < N) {
            cplx<double> t = lane == c ? cplx<double>{1, 0} : X[lane * N + c];
            X[lane * N + c] = t * inv;
        }
        wave_sync();
        // other rows: X(r, j) -= X(r, c) X(c, j), with X(r, c) := 0 first;
        // column c is snapshotted into the G area so no lane reads an
        // entry another lane has already rewritten (row c is not written)
        cplx<double>* col = reinterpret_cast<cplx<double>*>(l.G);
        if (lane < N) col[lane] = X[c * N + lane];
        wave_sync();
        for (int e = lane; e < N * N; e += 64) {
            int r = e % N, j = e / N;
            if (r == c) continue;
            cplx<double> x = j == c ? cplx<double>{0, 0} : X[e];
            X[e] = x - col[r] * X[j * N + c];
        }
        wave_sync();
    }
    // unscramble: columns swapped in reverse pivot order
    for (int c = N - 1; c >= 0; c--) {
        int p = __shfl(ipiv_lane, c);
        if (p != c) {
            if (lane < N) {
                cplx<double> a = X[p * N + lane], b = X[c * N + lane];
                X[p * N + lane] = b;
                X[c * N + lane] = a;
            }
            wave_sync();
        }
    }
    return ok;
}

// fresh FP64 inverse of A(sigma), antisymmetrised (pfaffian.cpp:284-297), into v
template <typename R>
__device__ inline bool pf_fresh(const PfLds<R>& l, const uint64_t* st, int lane, cplx<R> (&v)[16]) {
    pf_build(l, st, lane);
    int ip = 0;
    bool ok = pf_gj(l, lane, ip);
    const int i = lane & 31, h = lane >> 5, N = l.N;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        int j = 16 * h + q;
        cplx<R> r{0, 0};
        if (i < N && j < N) {
            cplx<double> a = l.Xd[j * N + i], b = l.Xd[i * N + j];
            r = {(R)(0.5 * (a.re - b.re)), (R)(0.5 * (a.im - b.im))};
        }
        v[q] = r;
    }
    wave_sync();
    return ok;
}

// log Pf(A(sigma)) by Parlett-Reid on X (pfaffian10 of pfaffian.cpp:284-297;
// own restatement); returns (-inf, 0) for a singular or odd matrix
template <typename R>
__device__ inline cplx<double> pf_logpf(const PfLds<R>& l, const uint64_t* st, int lane) {
    pf_build(l, st, lane);
    const int N = l.N;
    cplx<double>* X = l.Xd;
    if (N & 1) return {-INFINITY, 0};
    double lre = 0, lim = 0;
    for (int k = 0; k + 1 < N; k += 2) {
        double key = -1;
        if (lane > k && lane < N) {
            cplx<double> x = X[lane * N + k];  // X(k, lane)
            key = (double)x.re * x.re + (double)x.im * x.im;
        }
        int kp = wave_argmax(key, lane);
        if (kp != k + 1) {
            if (lane < N) {  // rows k+1 <-> kp
                cplx<double> a = X[lane * N + k + 1], b = X[lane * N + kp];
                X[lane * N + k + 1] = b;
                X[lane * N + kp] = a;
            }
            wave_sync();
            if (lane < N) {  // columns k+1 <-> kp
                cplx<double> a = X[(k + 1) * N + lane], b = X[kp * N + lane];
                X[(k + 1) * N + lane] = b;
                X[kp * N + lane] = a;
            }
            wave_sync();
            lim += M_PI;  // sign flip
        }
        cplx<double> a = X[(k + 1) * N + k];
        if (a.re == 0 && a.im == 0) return {-INFINITY, 0};
        lre += 0.5 * ::log((double)a.re * a.re + (double)a.im * a.im);
        lim += ::atan2((double)a.im, (double)a.re);
        int n2 = N - k - 2;
        for (int e = lane; e < n2 * n2; e += 64) {
            int i = k + 2 + e % n2, j = k + 2 + e / n2;
            cplx<double> up = X[i * N + k + 1] * X[j * N + k] - X[i * N + k] * X[j * N + k + 1];
            X[j * N + i] = X[j * N + i] + cdiv(up, a);
        }
        wave_sync();
    }
    return {lre, lim};
}

// Pfaffian of the 2m x 2m skew matrix G (row-major, stride 16): closed forms
// for m <= 2, Parlett-Reid by one lane otherwise (uniform result)
template <typename R>
__device__ inline cplx<R> small_pf(cplx<R>* G, int m, int lane) {
    if (m == 1) return G[1];
    if (m == 2) return G[1] * G[2 * 16 + 3] - G[2] * G[1 * 16 + 3] + G[3] * G[1 * 16 + 2];
    cplx<R> pf{1, 0};
    if (lane == 0) {
        const int n = 2 * m;
        for (int k = 0; k + 1 < n; k += 2) {
            int kp = k + 1;
            R best = norm2(G[k * 16 + k + 1]);
            for (int i = k + 2; i < n; i++)
                if (norm2(G[k * 16 + i]) > best) best = norm2(G[k * 16 + i]), kp = i;
            if (kp != k + 1) {
                for (int i = 0; i < n; i++) {
                    cplx<R> t = G[i * 16 + k + 1];
                    G[i * 16 + k + 1] = G[i * 16 + kp];
                    G[i * 16 + kp] = t;
                }
                for (int i = 0; i < n; i++) {
                    cplx<R> t = G[(k + 1) * 16 + i];
                    G[(k + 1) * 16 + i] = G[kp * 16 + i];
                    G[kp * 16 + i] = t;
                }
                pf = cplx<R>{-pf.re, -pf.im};
            }
            cplx<R> a = G[k * 16 + k + 1];
            pf = pf * a;
            if (a.re == 0 && a.im == 0) break;
            for (int i = k + 2; i < n; i++)
                for (int j = k + 2; j < n; j++)
                    G[i * 16 + j] = G[i * 16 + j] +
                                    cdiv(G[(k + 1) * 16 + i] * G[k * 16 + j] - G[k * 16 + i] * G[(k + 1) * 16 + j], a);
        }
    }
    return shfl_c(pf, 0);
}

// The Woodbury pieces of one flip set.  KEEP keeps the columns of inv B of
// this lane's row (m <= 2) for a following accept.
template <typename R, bool KEEP>
struct PfProp {
    cplx<R> ib[KEEP ? 4 : 1];
    int m;
};

// Pf(A(sigma')) / Pf(A(sigma)) for sigma' = sigma with the sites fl[i]
// (bit i of mask set, i < 8) flipped, in increasing i (uniform over the
// wave); leaves Cinv' in G (row-major, stride 16).  KEEP: at most the
// first two slots are used (single and bond flips) and the columns of
// inv B of this lane's row are kept for pf_accept.
template <typename R, bool KEEP>
__device__ inline cplx<R> pf_ratio(const PfLds<R>& l, const cplx<R> (&v)[16], const uint64_t* st, const int* fl,
                                   int mask, int lane, PfProp<R, KEEP>& pp) {
    constexpr int NS = KEEP ? 2 : 8;  // flip slots
    const int N = l.N, i = lane & 31, h = lane >> 5;
    const int m = __builtin_popcount(mask);
    pp.m = m;
    if constexpr (KEEP)
#pragma unroll
        for (int u = 0; u < 4; u++) pp.ib[u] = {0, 0};
    // compacted flip list (uniform) -> F
    if (lane == 0) {
        int k = 0;
#pragma unroll
        for (int t = 0; t < NS; t++)
            if ((mask >> t) & 1) l.F[k++] = fl[t];
    }
    wave_sync();
    int si = i < N ? bit2(st, i) : 0, fli = 0;
    for (int k = 0; k < m; k++) fli |= l.F[k] == i;
    // Delta-a rows: D[k][i] = a(i, f_k; flipped) - a(i, f_k)
    for (int k = 0; k < m; k++) {
        int f = l.F[k], sf = bit2(st, f);
        cplx<R> d{0, 0};
        if (i < N && i != f) d = pf_a(l, i, f, si ^ fli, sf ^ 1) - pf_a(l, i, f, si, sf);
        if (h == 0) l.D[k * 32 + i] = d;
    }
    wave_sync();
    // G = B^T inv B: column b < m is inv Delta-a_b, column m + b is inv e_{f_b}
    for (int b = 0; b < 2 * m; b++) {
        cplx<R> col;
        if (b < m) {
            cplx<R> s{0, 0};
#pragma unroll
            for (int q = 0; q < 16; q++) s = s + v[q] * l.D[b * 32 + 16 * h + q];
            col = {s.re + partner<32>(s.re), s.im + partner<32>(s.im)};
        } else {
            int f = l.F[b - m];
            cplx<R> sel{0, 0};
#pragma unroll
            for (int q = 0; q < 16; q++)
                if (q == (f & 15)) sel = v[q];
            cplx<R> oth{partner<32>(sel.re), partner<32>(sel.im)};
            col = (f >> 4) == h ? sel : oth;
        }
        if constexpr (KEEP) {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (u == b) pp.ib[u] = col;
        }
        for (int a = 0; a < m; a++) {
            // rows a < m: sum_i Delta-a_a(i) col(i); rows m + a: col(f_a)
            cplx<R> x = group_csum<32>(l.D[a * 32 + i] * col);
            int fa = __builtin_amdgcn_readfirstlane(l.F[a]);
            cplx<R> r{__shfl(col.re, fa), __shfl(col.im, fa)};
            if (lane == 0) {
                l.G[a * 16 + b] = x;
                l.G[(m + a) * 16 + b] = r;
            }
        }
    }
    wave_sync();
    // Cinv' = Cinv0 + (G - G^T) / 2: unit entries and the antisymmetric part
    const int n = 2 * m, nE = n * n;
    constexpr int NU = (2 * NS * 2 * NS + 63) / 64;
    cplx<R> g[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) {
        int e = lane + 64 * u;
        cplx<R> val{0, 0};
        if (e < nE) {
            int a = e / n, b = e % n;
            cplx<R> gab = l.G[a * 16 + b], gba = l.G[b * 16 + a];
            val = {(R)0.5 * (gab.re - gba.re), (R)0.5 * (gab.im - gba.im)};
            if (a >= m && b == a - m) val.re += 1;
            if (b >= m && a == b - m) val.re -= 1;
        }
        g[u] = val;
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < NU; u++) {
        int e = lane + 64 * u;
        if (e < nE) l.G[(e / n) * 16 + e % n] = g[u];
    }
    wave_sync();
    // pair entries Cinv0(k, q) = a(f_k, f_q) - a(f_k, f_q; both flipped), k > q
    if (lane == 0)
        for (int k = 1; k < m; k++)
            for (int q = 0; q < k; q++) {
                int fk = l.F[k], fq = l.F[q], sk = bit2(st, fk), sq = bit2(st, fq);
                cplx<R> c0 = pf_a(l, fk, fq, sk, sq) - pf_a(l, fk, fq, sk ^ 1, sq ^ 1);
                l.G[k * 16 + q] = l.G[k * 16 + q] + c0;
                l.G[q * 16 + k] = l.G[q * 16 + k] - c0;
            }
    wave_sync();
    cplx<R> pf = small_pf(l.G, m, lane);
    if (m > 2) wave_sync();
    bool neg = ((m * (m + 1) / 2) & 1) != 0;
    return neg ? cplx<R>{-pf.re, -pf.im} : pf;
}

// accept the flip set of the last pf_ratio<R, true> (m <= 2), whose Cinv'
// is in G.  C = antisym(Cinv'^-1 / 2) = Cinv'^-1 for the skew Cinv', by
// the closed forms of the 2 x 2 and 4 x 4 skew inverses:
//   [[0, x], [-x, 0]]^-1 = [[0, -1/x], [1/x, 0]],
//   A^-1 = (1/Pf A) [[0, -a23, a13, -a12], [a23, 0, -a03, a02],
//                    [-a13, a03, 0, -a01], [a12, -a02, a01, 0]].
template <typename R>
__device__ inline void pf_accept(const PfLds<R>& l, cplx<R> (&v)[16], const PfProp<R, true>& pp, int lane) {
    const int m = pp.m, n = 2 * m, i = lane & 31, h = lane >> 5;
    // upper triangle c01, c02, c03, c12, c13, c23 of C
    cplx<R> c01, c02{0, 0}, c03{0, 0}, c12{0, 0}, c13{0, 0}, c23{0, 0};
    if (m == 1) {
        c01 = cdiv(cplx<R>{-1, 0}, l.G[1]);
    } else {
        cplx<R> a01 = l.G[1], a02 = l.G[2], a03 = l.G[3], a12 = l.G[16 + 2], a13 = l.G[16 + 3],
                a23 = l.G[32 + 3];
        cplx<R> ip = cdiv(cplx<R>{1, 0}, a01 * a23 - a02 * a13 + a03 * a12);
        c01 = cplx<R>{-a23.re, -a23.im} * ip;
        c02 = a13 * ip;
        c03 = cplx<R>{-a12.re, -a12.im} * ip;
        c12 = cplx<R>{-a03.re, -a03.im} * ip;
        c13 = a02 * ip;
        c23 = cplx<R>{-a01.re, -a01.im} * ip;
    }
    auto neg = [](cplx<R> x) { return cplx<R>{-x.re, -x.im}; };
    // t1 = (inv B) C for this lane's row
    cplx<R> t1[4];
    t1[0] = neg(pp.ib[1] * c01) - pp.ib[2] * c02 - pp.ib[3] * c03;
    t1[1] = pp.ib[0] * c01 - pp.ib[2] * c12 - pp.ib[3] * c13;
    t1[2] = pp.ib[0] * c02 + pp.ib[1] * c12 - pp.ib[3] * c23;
    t1[3] = pp.ib[0] * c03 + pp.ib[1] * c13 + pp.ib[2] * c23;
    // exchange rows of inv B and t1 through X: X[b*32 + i], X[(4 + b)*32 + i]
    if (h == 0)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            if (b < n) {
                l.X[b * 32 + i] = pp.ib[b];
                l.X[(4 + b) * 32 + i] = t1[b];
            }
        }
    wave_sync();
    // inv(i, j) += (1/2) sum_b [t1(i, b) invB(j, b) - t1(j, b) invB(i, b)]
#pragma unroll
    for (int q = 0; q < 16; q++) {
        int j = 16 * h + q;
        cplx<R> s{0, 0};
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (b < n) s = s + t1[b] * l.X[b * 32 + j] - l.X[(4 + b) * 32 + j] * pp.ib[b];
        v[q] = v[q] + cplx<R>{(R)0.5 * s.re, (R)0.5 * s.im};
    }
    wave_sync();
}

// d ln Pf / d fs (pfaffian.cpp:399-415) summed into out[n_pf] (this wave's
// accumulators in X, then written by the lanes)
template <typename R>
__device__ inline void pf_derivative(const PfLds<R>& l, const cplx<R> (&v)[16], const uint64_t* st, int lane,
                                     cplx<R>* out) {
    const int N = l.N, i = lane & 31, h = lane >> 5, npf = 4 * l.rows;
    for (int e = lane; e < npf; e += 64) l.X[e] = {0, 0};
    wave_sync();
    if (i < N) {
        int si = bit2(st, i);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            int j = 16 * h + q;
            if (j < i) {
                int bb = l.bs[i * N + j];
                int idx = (abs(bb) - 1) + ((si << 1) | bit2(st, j)) * l.rows;
                cplx<R> x = v[q];
                if (bb > 0) x = {-x.re, -x.im};
                atomicAdd(&l.X[idx].re, x.re);
                atomicAdd(&l.X[idx].im, x.im);
            }
        }
    }
    wave_sync();
    for (int e = lane; e < npf; e += 64) out[e] = l.X[e];
    wave_sync();
}

}  // namespace rbmgpu
