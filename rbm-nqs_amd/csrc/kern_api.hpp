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
    int pf_refresh;  // fresh Pfaffian inverse every pf_refresh accepted moves (0: never)
    const int32_t* pb;
    uint64_t* states;
    double* obs;
    cplx<R>* Tout;
    cplx<R>* Opf;    // S x 4 pf_rows: d ln Pf / d fs per kept sample
    uint32_t* acc;
    uint32_t* prop;
};

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
    double* part;
};

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
    cplx<R>* part;       // nblocks x ld
    double* part2;       // nchunks x P complex
    double* y;           // complex P
    double* pdot;        // per column block: p^H y partial (2 doubles), optional
    const CgStatus* status;
};

// structured (basic RBM) S.x / moments, FP32 MFMA; see kern_struct.hpp
struct StructArgs {
    int64_t S;
    int N, M, P, ld, nw, n_groups, mode;  // mode 0: S.x, 1: y = E_loc, 2: y = 1
    const uint64_t* states;
    const cplx<float>* T;
    const double* obs;
    const double* x;
    cplx<float>* part;  // nblocks x ld: [z_v | z_h | Z | sum |t|^2]
    const CgStatus* status;
};
// LDS bytes of k_struct for padded sizes, 0 if no instance covers N, M
size_t struct_lds_bytes(int N, int M);
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

// supported (lanes per chain, units per lane) shapes
struct Shape {
    int L, K;
};
constexpr Shape kShapes[] = {{8, 1},  {8, 2},  {8, 4},  {8, 8},  {8, 16}, {16, 2},
                             {16, 4}, {16, 8}, {16, 16}, {32, 2}, {32, 4}, {32, 8},
                             {32, 16}, {64, 1}, {64, 2}, {64, 4}, {64, 8}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <typename R>
size_t lds_bytes(const Dims& d) {
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

}  // namespace rbmgpu
