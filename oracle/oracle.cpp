/*
 * oracle.cpp -- FP64 CPU oracle of the reference's VMC path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function restates the
 * reference file:line named in its comment.  It is compiled with g++ by
 * oracle/build.sh into oracle/_build/liboracle.so.
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   lattice      -- honeycomb(3) neighbour KAT (test_lattice.cpp:242-255), the
 *                   bonds3 / plaqs3 lists of notebooks/KitaevED.ipynb cell 6,
 *                   bravais up/down wrap tables (test_lattice.cpp:53-223);
 *   Hamiltonian  -- ED energies -0.8660254037844461 (2x2) and
 *                   -0.7939723678960895 (3x3) (KitaevED.ipynb cells 10, 13),
 *                   the toric-code ground energy -1, kitaev.cpp:66-100;
 *   RBM          -- incremental vs fresh theta (test_derivative.cpp:53-76),
 *                   Taylor derivative check (78-112), e^{dlogpsi} == ratio,
 *                   symmetry invariance psi(P_s sigma) == psi(sigma);
 *   Pfaffian     -- Pf^2 == det, 4x4 closed form, SM update vs fresh
 *                   (test_derivative.cpp:120-154), Taylor check (156-195);
 *   aggregator   -- closed-form mean / unbiased variance / binned tau;
 *   SR           -- S.x vs explicit S, CG residual;
 *   sampler      -- stationary distribution vs exact enumeration.
 */
#include "oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using cd = std::complex<double>;
static thread_local std::string g_err;

/* ------------------------------------------------------------------ RNG --
 * Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the published
 * multiplier and Weyl constants.  Streams: key = seed, counter =
 * (chain, epoch, step, slot). */
static inline void philox(const uint32_t c[4], const uint32_t k[2], uint32_t o[4]) {
    uint32_t x0 = c[0], x1 = c[1], x2 = c[2], x3 = c[3], k0 = k[0], k1 = k[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o[0] = x0; o[1] = x1; o[2] = x2; o[3] = x3;
}
static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) *
           (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------ ini --
 * Subset of ini.cpp:38-122 keys, same names and defaults. */
struct Decay {
    double initial, min, decay;
};
struct Config {
    std::string model = "kitaev";
    size_t n_cells = 2;
    int n_cells_b = -1;
    std::vector<double> J{-1.};
    double h = 0, helper = 0;
    std::vector<double> symmetry{0.5};
    std::string rbm = "basic";
    size_t alpha = 1;
    bool pfaffian = false;
    std::vector<double> pf_symmetry{0.5};
    size_t n_samples = 1000, spc = 0, n_chains = 4, warmup = 100, step = 1;
    double bond_flips = 0.5;
    int pf_refresh = 0;
    Decay lr{1e-2, 1e-2, 1}, reg1{1, 1e-4, 0.9}, reg2{1e-3, 1e-6, 0.9},
        dreg1{1e-2, 1e-2, 1};
    size_t sr_max_iter = 0;
    double sr_rtol = 0;
};

static std::vector<double> parse_list(const std::string& s) {
    std::vector<double> r;
    std::string t;
    for (char c : s) {
        if (c == ',' || c == ' ' || c == '\t') {
            if (!t.empty()) r.push_back(std::stod(t));
            t.clear();
        } else
            t += c;
    }
    if (!t.empty()) r.push_back(std::stod(t));
    return r;
}
static Decay parse_decay(const std::string& s) {
    auto v = parse_list(s);
    if (v.size() == 3) return {v[0], v[1], v[2]};
    return {v.at(0), v.at(0), 1.};
}
static bool parse_bool(const std::string& s) {
    return s == "1" || s == "true" || s == "True" || s == "yes" || s == "on";
}

static Config parse_ini(const std::string& text) {
    Config c;
    std::istringstream in(text);
    std::string line, section;
    while (std::getline(in, line)) {
        auto p = line.find_first_of("#;");
        if (p != std::string::npos) line = line.substr(0, p);
        auto trim = [](std::string s) {
            size_t a = s.find_first_not_of(" \t\r\n");
            size_t b = s.find_last_not_of(" \t\r\n");
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        line = trim(line);
        if (line.empty()) continue;
        if (line[0] == '[') {
            section = trim(line.substr(1, line.find(']') - 1));
            continue;
        }
        auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (!section.empty()) key = section + "." + key;
        if (key == "model.type") c.model = val;
        else if (key == "model.n_cells") c.n_cells = std::stoul(val);
        else if (key == "model.n_cells_b") c.n_cells_b = std::stoi(val);
        else if (key == "model.J") c.J = parse_list(val);
        else if (key == "model.h") c.h = std::stod(val);
        else if (key == "model.helper_strength") c.helper = std::stod(val);
        else if (key == "model.symmetry") c.symmetry = parse_list(val);
        else if (key == "rbm.type") c.rbm = val;
        else if (key == "rbm.alpha") c.alpha = std::stoul(val);
        else if (key == "rbm.pfaffian") c.pfaffian = parse_bool(val);
        else if (key == "rbm.pfaffian.symmetry") c.pf_symmetry = parse_list(val);
        else if (key == "sampler.n_samples") c.n_samples = std::stoul(val);
        else if (key == "sampler.n_samples_per_chain") c.spc = std::stoul(val);
        else if (key == "sampler.metropolis.n_chains") c.n_chains = std::stoul(val);
        else if (key == "sampler.metropolis.n_warmup_steps") c.warmup = std::stoul(val);
        else if (key == "sampler.metropolis.n_steps_per_sample") c.step = std::stoul(val);
        else if (key == "sampler.metropolis.bond_flips") c.bond_flips = std::stod(val);
        else if (key == "sampler.pfaffian_refresh") c.pf_refresh = std::stoi(val);
        else if (key == "optimizer.learning_rate") c.lr = parse_decay(val);
        else if (key == "optimizer.sr.reg1") c.reg1 = parse_decay(val);
        else if (key == "optimizer.sr.reg2") c.reg2 = parse_decay(val);
        else if (key == "optimizer.sr.deltareg1") c.dreg1 = parse_decay(val);
        else if (key == "optimizer.sr.max_iterations") c.sr_max_iter = std::stoul(val);
        else if (key == "optimizer.sr.rtol") c.sr_rtol = std::stod(val);
    }
    if (c.spc) c.n_samples = c.n_chains * c.spc; /* ini.cpp:259-262 */
    return c;
}

/* -------------------------------------------------------------- lattice --
 * bravais.cpp:26-284, honeycomb.cpp:30-209, toric_lattice.cpp:30-69. */
struct Lattice {
    int kind = 0; /* 0 honeycomb, 1 toric */
    size_t n_uc, n_dim = 2, n_basis = 2, n_coord, n_total_uc, n_total, n_uc_b,
                 h_shift = 0;
    std::vector<double> default_symm;
    std::vector<std::array<size_t, 3>> bonds;

    size_t uc_idx(size_t i) const { return i / n_basis; }
    size_t b_idx(size_t i) const { return i % n_basis; }
    size_t idx(size_t uc, size_t b) const { return uc * n_basis + b; }

    size_t up(size_t uc, size_t dir = 0, size_t step = 1) const { /* bravais.cpp:66-85 */
        if (step == 0) return uc;
        size_t shift = (size_t)std::pow((double)n_uc, (double)dir) * n_uc_b;
        size_t offset = uc / shift;
        size_t n2 = dir == 0 ? n_uc_b : n_uc;
        size_t overflow = (uc + (step * shift) / n2) / shift;
        step %= n2;
        uc = (uc + step * shift / n2) % shift;
        size_t r = shift * offset + uc;
        if (overflow && h_shift && dir == 1) r = down(r, 0, h_shift);
        return r;
    }
    size_t down(size_t uc, size_t dir = 0, size_t step = 1) const { /* bravais.cpp:86-105 */
        if (step == 0) return uc;
        size_t shift = (size_t)std::pow((double)n_uc, (double)dir) * n_uc_b;
        size_t offset = uc / shift;
        size_t n2 = dir == 0 ? n_uc_b : n_uc;
        size_t overflow = (uc + (n2 - step) * shift / n2) / shift;
        step %= n2;
        uc = (uc + (n2 - step) * shift / n2) % shift;
        size_t r = shift * offset + uc;
        if (!overflow && h_shift && dir == 1) r = up(r, 0, h_shift);
        return r;
    }
    std::vector<size_t> nns(size_t i) const {
        size_t uc = uc_idx(i);
        if (kind == 0) { /* honeycomb.cpp:36-45 */
            if (b_idx(i) == 0) return {idx(uc, 1), idx(down(uc), 1), idx(down(uc, 1), 1)};
            return {idx(uc, 0), idx(up(uc), 0), idx(up(uc, 1), 0)};
        }
        /* toric_lattice.cpp:34-45 */
        if (b_idx(i) == 0)
            return {idx(uc, 1), idx(up(uc), 1), idx(down(uc, 1), 1), idx(up(down(uc, 1)), 1)};
        return {idx(uc, 0), idx(down(uc), 0), idx(up(uc, 1), 0), idx(down(up(uc, 1)), 0)};
    }
    void construct_bonds() { /* honeycomb.cpp:47-55, toric_lattice.cpp:47-55 */
        for (size_t i = 0; i < n_total_uc; i++) {
            auto nn = nns(idx(i, 0));
            for (size_t c = 0; c < n_coord; c++) bonds.push_back({idx(i, 0), nn[c], c});
        }
    }
    std::vector<std::vector<size_t>> hexagons() const { /* honeycomb.cpp:201-209 */
        std::vector<std::vector<size_t>> r;
        for (size_t uc = 0; uc < n_total_uc; uc++)
            r.push_back({idx(down(uc, 1), 1), idx(uc, 0), idx(uc, 1), idx(up(uc), 0),
                         idx(down(up(uc), 1), 1), idx(down(up(uc), 1), 0)});
        return r;
    }
    std::vector<std::pair<std::array<size_t, 4>, size_t>> plaqs() const { /* toric_lattice.cpp:57-69 */
        std::vector<std::pair<std::array<size_t, 4>, size_t>> r;
        for (size_t uc = 0; uc < n_total_uc; uc++) {
            r.push_back({{idx(uc, 0), idx(uc, 1), idx(up(uc, 1), 0), idx(up(uc), 1)}, 0});
            r.push_back({{idx(uc, 0), idx(uc, 1), idx(down(uc), 0), idx(down(uc, 1), 1)}, 1});
        }
        return r;
    }
    std::vector<size_t> clean_symms(const std::vector<double>& symm) const { /* bravais.cpp:249-284 */
        int factor = 1;
        for (double s : symm) {
            if (s < 0) throw std::runtime_error("Negative Symmetry factors are not allowed!");
            if (std::abs(factor * s - int(factor * s)) > 1e-14) {
                double r = factor * s - std::floor(factor * s);
                factor *= (int)(1 / r);
            }
        }
        std::vector<size_t> x(n_dim);
        if (symm.size() != 1 && symm.size() != n_dim)
            throw std::runtime_error("Symmetry factor sizes incompatible with lattice dimensions.");
        for (size_t i = 0; i < symm.size(); i++) {
            x[i] = int(factor * symm[i]);
            if (x[i] == 0) x[i] = (i == 0 ? n_uc_b : n_uc);
            if ((i == 0 && n_uc_b % x[i] != 0) || (i > 0 && n_uc % x[i] != 0))
                throw std::runtime_error("Symmetry factor invalid");
        }
        if (symm.size() == 1)
            for (size_t i = 1; i < n_dim; i++) x[i] = x[0];
        return x;
    }
    static bool is_half(const std::vector<double>& s) { return s.size() == 1 && s[0] == 0.5; }
    static bool is_none(const std::vector<double>& s) { return s.empty() || s[0] == 0; }

    std::vector<std::vector<size_t>> uc_symmetry(const std::vector<double>& symm) const {
        /* bravais.cpp:119-160 */
        if (is_none(symm)) {
            std::vector<size_t> r(n_total_uc);
            for (size_t i = 0; i < n_total_uc; i++) r[i] = i;
            return {r};
        }
        if (kind == 0 && symm.size() == 1 && std::abs(symm[0] - 0.6) < 1e-10)
            throw std::runtime_error("hex symmetry 0.6 not supported");
        auto x = clean_symms(symm);
        std::vector<std::vector<size_t>> ret;
        std::vector<size_t> shifts(n_dim);
        std::function<void(size_t)> all = [&](size_t d) {
            if (d >= n_dim) {
                std::vector<size_t> r(n_total_uc);
                for (size_t u = 0; u < n_total_uc; u++) {
                    size_t uc = u;
                    for (size_t i = 0; i < shifts.size(); i++) uc = up(uc, i, shifts[i]);
                    r[u] = uc;
                }
                ret.push_back(r);
            } else {
                for (size_t i = 0; i < (d == 0 ? n_uc_b : n_uc); i += x[d]) {
                    shifts[d] = i;
                    all(d - 1);
                }
            }
        };
        all(n_dim - 1);
        return ret;
    }
    /* perms[s][i] = indices()(i) of the s-th Eigen permutation */
    std::vector<std::vector<size_t>> symmetry(const std::vector<double>& symm) const {
        if (is_none(symm)) {
            std::vector<size_t> r(n_total);
            for (size_t i = 0; i < n_total; i++) r[i] = i;
            return {r};
        }
        auto ucs = uc_symmetry(symm);
        if (kind == 0 && is_half(symm)) { /* honeycomb.cpp:84-130 */
            std::vector<std::vector<size_t>> ret(n_total, std::vector<size_t>(n_total));
            for (size_t i = 0; i < n_total_uc; i++)
                for (size_t k = 0; k < n_total; k++) {
                    size_t v = idx(ucs[i][uc_idx(k)], b_idx(k));
                    ret[2 * i][k] = v;
                    ret[2 * i + 1][k] = n_total - 1 - v; /* rot180 */
                }
            return ret;
        }
        std::vector<std::vector<size_t>> ret(ucs.size(), std::vector<size_t>(n_total));
        for (size_t s = 0; s < ucs.size(); s++) /* bravais.cpp:162-205 */
            for (size_t k = 0; k < n_total; k++) ret[s][k] = idx(ucs[s][uc_idx(k)], b_idx(k));
        return ret;
    }
    std::vector<size_t> symm_basis(const std::vector<double>& symm) const {
        if (kind == 0 && is_half(symm)) return {0}; /* honeycomb.cpp:136-145 */
        if (is_none(symm)) {                           /* bravais.cpp:207-235 */
            std::vector<size_t> r(n_total);
            for (size_t i = 0; i < n_total; i++) r[i] = i;
            return r;
        }
        auto x = clean_symms(symm);
        std::vector<size_t> ucs;
        std::function<void(size_t, size_t)> get = [&](size_t d, size_t uc) {
            if (d >= n_uc) ucs.push_back(uc);
            else
                for (size_t i = 0; i < x[d]; i++) get(d - 1, up(uc, d, i));
        };
        get(n_dim - 1, 0);
        std::vector<size_t> r;
        for (auto uc : ucs)
            for (size_t b = 0; b < n_basis; b++) r.push_back(idx(uc, b));
        return r;
    }
    size_t symmetry_size(const std::vector<double>& symm) const {
        if (is_none(symm)) return 1;
        if (kind == 0 && is_half(symm)) return n_total;
        auto x = clean_symms(symm);
        size_t r = n_uc_b / x[0];
        for (size_t i = 1; i < n_dim; i++) r *= n_uc / x[i];
        return r;
    }
};

static Lattice make_honeycomb(size_t n, int nb, const std::vector<double>& symm) {
    Lattice l;
    l.kind = 0;
    l.n_uc = n;
    l.n_coord = 3;
    l.n_uc_b = nb == -1 ? n : (size_t)nb;
    l.n_total_uc = n * l.n_uc_b;
    l.n_total = l.n_total_uc * 2;
    l.default_symm = symm;
    l.construct_bonds();
    return l;
}
static Lattice make_toric(size_t n) {
    Lattice l;
    l.kind = 1;
    l.n_uc = n;
    l.n_coord = 4;
    l.n_uc_b = n;
    l.n_total_uc = n * n;
    l.n_total = 2 * n * n;
    l.default_symm = {1};
    l.construct_bonds();
    return l;
}

/* --------------------------------------------------------------- model --
 * kitaev.cpp:29-64, toric.cpp:32-47, abstract_model.cpp:28-54.  A local op
 * is a dense 2^k x 2^k matrix; acts_on[0] is the least significant bit of
 * the local index (kron ordering, abstract_model.cpp:31). */
struct LocalOp {
    std::vector<size_t> sites;
    size_t dim;
    std::vector<cd> mat; /* row-major dim x dim */
};
struct Mat {
    size_t n;
    std::vector<cd> a;
};
static Mat kron_list(const std::vector<Mat>& args) {
    Mat so{1, {cd(1)}};
    for (const auto& A : args) { /* so = A (x) so */
        Mat r{A.n * so.n, std::vector<cd>(A.n * so.n * A.n * so.n)};
        for (size_t i = 0; i < A.n; i++)
            for (size_t j = 0; j < A.n; j++)
                for (size_t k = 0; k < so.n; k++)
                    for (size_t l = 0; l < so.n; l++)
                        r.a[(i * so.n + k) * r.n + j * so.n + l] = A.a[i * A.n + j] * so.a[k * so.n + l];
        so = r;
    }
    return so;
}
static Mat sx() { return {2, {0., 1., 1., 0.}}; }
static Mat sy() { return {2, {0., cd(0, -1), cd(0, 1), 0.}}; }
static Mat sz() { return {2, {1., 0., 0., -1.}}; }
static Mat scale(Mat m, cd s) {
    for (auto& x : m.a) x *= s;
    return m;
}

struct Model {
    Config cfg;
    Lattice lat;
    std::vector<LocalOp> ham;
    std::array<double, 3> J{-1, -1, -1};
    double exact_energy() const { /* kitaev.cpp:66-100 */
        if (lat.kind != 0) return cfg.J.at(0);
        double smret = -1e10, na = lat.n_uc_b, nb = lat.n_uc;
        for (int c1i = 0; c1i < 2; c1i++)
            for (int c2i = 0; c2i < 2; c2i++) {
                double ret = 0, c1 = 0.5 / nb * c1i, c2 = 0.5 / na * c2i;
                for (size_t i = 0; i < nb; i++)
                    for (size_t j = 0; j < na; j++) {
                        double fa = i / nb, fb = j / na;
                        cd f = 2. * (J[0] * std::exp(cd(0, 2 * M_PI * (fa + c1))) +
                                     J[1] * std::exp(cd(0, 2 * M_PI * (fb + c2))) + J[2]);
                        ret += std::abs(f);
                    }
                if (ret > smret) smret = ret;
            }
        return -0.5 * smret / lat.n_total;
    }
};

static void add_op(Model& m, const std::vector<size_t>& sites, const Mat& op) {
    m.ham.push_back({sites, op.n, op.a});
}

static void build_model(Model& m) {
    const Config& c = m.cfg;
    if (c.model == "kitaev") {
        m.lat = make_honeycomb(c.n_cells, c.n_cells_b, c.symmetry);
        for (int i = 0; i < 3; i++) m.J[i] = c.J[i % c.J.size()];
        Mat ops[3] = {scale(kron_list({sx(), sx()}), m.J[0]), scale(kron_list({sy(), sy()}), m.J[1]),
                      scale(kron_list({sz(), sz()}), m.J[2])};
        for (auto& b : m.lat.bonds) add_op(m, {b[0], b[1]}, ops[b[2]]);
        if (std::abs(c.h) > 1e-10) {
            Mat op{2, std::vector<cd>(4)};
            Mat x = sx(), y = sy(), z = sz();
            for (int i = 0; i < 4; i++) op.a[i] = -c.h * x.a[i] + c.h * y.a[i] + c.h * z.a[i];
            for (size_t i = 0; i < m.lat.n_total; i++) add_op(m, {i}, op);
        }
        if (c.helper != 0.) { /* kitaev.cpp:52-64 */
            Mat p = kron_list({sx(), sy(), sz(), sx(), sy(), sz()});
            for (size_t i = 0; i < p.n; i++) p.a[i * p.n + i] -= 1.;
            p = scale(p, -c.helper);
            for (auto& hx : m.lat.hexagons()) add_op(m, hx, p);
        }
    } else if (c.model == "toric") {
        m.lat = make_toric(c.n_cells);
        double J = c.J.at(0);
        Mat plaq = scale(kron_list({sz(), sz(), sz(), sz()}), J);
        Mat vert = scale(kron_list({sx(), sx(), sx(), sx()}), J);
        for (auto& p : m.lat.plaqs())
            add_op(m, {p.first[0], p.first[1], p.first[2], p.first[3]}, p.second == 0 ? plaq : vert);
    } else {
        throw std::runtime_error("Model Type '" + c.model + "' not available!");
    }
}

/* ---------------------------------------------------------- spin state --
 * spin_state.cpp:44-82. */
struct State {
    std::array<uint64_t, 2> w{0, 0};
    bool operator[](size_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
    void flip(size_t i) { w[i >> 6] ^= (uint64_t)1 << (i & 63); }
    void flip(const std::vector<size_t>& f) {
        for (auto i : f) flip(i);
    }
    double sig(size_t i) const { return (*this)[i] ? -1. : 1.; }
    bool operator==(const State& o) const { return w == o.w; }
    size_t to_num_loc(const std::vector<size_t>& acts) const {
        size_t r = 0;
        for (size_t i = 0; i < acts.size(); i++)
            if ((*this)[acts[i]]) r |= (size_t)1 << i;
        return r;
    }
};
struct StateHash {
    size_t operator()(const State& s) const { return std::hash<uint64_t>()(s.w[0] * 0x9E3779B97F4A7C15ull ^ s.w[1]); }
};
static State load_state(const uint64_t* p, size_t nw) {
    State s;
    for (size_t i = 0; i < nw; i++) s.w[i] = p[i];
    return s;
}

/* ------------------------------------------------------------ pfaffian --
 * Dense helpers: complex matrices are column-major n x n here, like Eigen. */
using MatC = std::vector<cd>;

/* Pfaffian of a skew-symmetric n x n matrix (column-major) by Parlett-Reid
 * reduction with pivoting (M. Wimmer, ACM TOMS 38 (2012) 30); own
 * derivation from Pf([[P,B],[-B^T,C]]) = Pf(P) Pf(C + B^T P^-1 B) with a 2x2
 * pivot block P.  Result as mantissa * 10^exp (the skpf10 convention). */
static cd pfaffian10(MatC A, size_t n, int& exp10) {
    exp10 = 0;
    if (n % 2) return 0.;
    cd pf = 1.;
    auto at = [&](size_t i, size_t j) -> cd& { return A[i + j * n]; };
    for (size_t k = 0; k + 1 < n; k += 2) {
        size_t kp = k + 1;
        double best = std::abs(at(k, k + 1));
        for (size_t i = k + 2; i < n; i++)
            if (std::abs(at(k, i)) > best) best = std::abs(at(k, i)), kp = i;
        if (kp != k + 1) {
            for (size_t i = 0; i < n; i++) std::swap(at(i, k + 1), at(i, kp));
            for (size_t i = 0; i < n; i++) std::swap(at(k + 1, i), at(kp, i));
            pf = -pf;
        }
        cd a = at(k, k + 1);
        if (a == 0.) {
            exp10 = 0;
            return 0.;
        }
        pf *= a;
        double lg = std::floor(std::log10(std::abs(pf)));
        if (std::isfinite(lg) && lg != 0) {
            pf *= std::pow(10., -lg);
            exp10 += (int)lg;
        }
        for (size_t i = k + 2; i < n; i++)
            for (size_t j = k + 2; j < n; j++)
                at(i, j) += (at(k + 1, i) * at(k, j) - at(k, i) * at(k + 1, j)) / a;
    }
    return pf;
}

/* inverse by Gauss-Jordan with partial pivoting (column-major) */
static bool inverse(MatC A, size_t n, MatC& inv) {
    inv.assign(n * n, 0.);
    for (size_t i = 0; i < n; i++) inv[i + i * n] = 1.;
    for (size_t c = 0; c < n; c++) {
        size_t p = c;
        for (size_t r = c + 1; r < n; r++)
            if (std::abs(A[r + c * n]) > std::abs(A[p + c * n])) p = r;
        if (A[p + c * n] == 0.) return false;
        if (p != c)
            for (size_t j = 0; j < n; j++) {
                std::swap(A[c + j * n], A[p + j * n]);
                std::swap(inv[c + j * n], inv[p + j * n]);
            }
        cd d = 1. / A[c + c * n];
        for (size_t j = 0; j < n; j++) A[c + j * n] *= d, inv[c + j * n] *= d;
        for (size_t r = 0; r < n; r++) {
            if (r == c) continue;
            cd f = A[r + c * n];
            if (f == 0.) continue;
            for (size_t j = 0; j < n; j++) {
                A[r + j * n] -= f * A[c + j * n];
                inv[r + j * n] -= f * inv[c + j * n];
            }
        }
    }
    return true;
}

struct PfCtx {
    MatC inv;
    cd pfaff = 1., update_factor = 1.;
    int exp = 0;
};

/* pfaffian.cpp:31-137 (index map and matrix elements), 270-435 (context,
 * Sherman-Morrison/Woodbury update, derivative, weight update) */
struct Pfaffian {
    size_t ns, rows;
    std::vector<int> bs; /* ns x ns, column-major like Eigen MatrixXi */
    MatC fs;             /* rows x 4, column-major */

    int& b(size_t i, size_t j) { return bs[i + j * ns]; }
    int b(size_t i, size_t j) const { return bs[i + j * ns]; }

    void build(const Lattice& lat, const std::vector<double>& symm) {
        ns = lat.n_total;
        auto perms = lat.symmetry(symm);
        auto basis = lat.symm_basis(symm);
        bs.assign(ns * ns, 0);
        int offset = 0;
        for (auto bi : basis) {
            for (auto& idcs : perms) {
                int f = idcs[0] > idcs[1] ? -1 : 1;
                for (size_t j = 0; j < ns; j++) {
                    int a = (int)idcs[bi], bb = (int)idcs[j];
                    b(bb, a) = (bb == a) ? 0 : f * (int)(j + offset);
                }
            }
            offset += (int)ns;
        }
        for (size_t i = 0; i < ns; i++)
            for (size_t j = 0; j < i; j++) {
                if (std::abs(b(i, j)) <= std::abs(b(j, i))) b(j, i) = -b(i, j);
                else b(i, j) = -b(j, i);
            }
        auto maxc = [&]() { return *std::max_element(bs.begin(), bs.end()); };
        for (int i = 1; i < maxc(); i++) {
            while (std::count(bs.begin(), bs.end(), i) == 0) {
                for (auto& x : bs) {
                    if (x > i) x -= 1;
                    else if (x < -i) x += 1;
                }
            }
        }
        rows = (size_t)maxc();
        fs.assign(rows * 4, 0.);
    }
    size_t n_params() const { return rows * 4; }
    int spidx(size_t i, size_t j, const State& s, bool fi = false, bool fj = false) const {
        return ((s[i] ^ fi) << 1) + (s[j] ^ fj);
    }
    cd a(size_t i, size_t j, const State& s, bool fi = false, bool fj = false) const {
        if (i == j) return 0.;
        int bb = b(i, j);
        double sgn = bb < 0 ? -1. : 1.;
        int sp = (j < i) ? spidx(i, j, s, fi, fj) : spidx(j, i, s, fj, fi);
        return sgn * fs[(size_t)(std::abs(bb) - 1) + (size_t)sp * rows];
    }
    MatC get_mat(const State& s) const {
        MatC m(ns * ns, 0.);
        for (size_t i = 0; i < ns; i++)
            for (size_t j = 0; j < i; j++) {
                m[i + j * ns] = a(i, j, s);
                m[j + i * ns] = -m[i + j * ns];
            }
        return m;
    }
    PfCtx get_context(const State& s) const {
        PfCtx c;
        MatC m = get_mat(s), inv;
        if (!inverse(m, ns, inv)) throw std::runtime_error("singular Pfaffian matrix");
        c.inv.assign(ns * ns, 0.);
        for (size_t i = 0; i < ns; i++)
            for (size_t j = 0; j < ns; j++) c.inv[i + j * ns] = 0.5 * (inv[i + j * ns] - inv[j + i * ns]);
        c.pfaff = pfaffian10(m, ns, c.exp);
        return c;
    }
    void update_context(const State& s, const std::vector<size_t>& flips, PfCtx& c) const {
        if (flips.empty()) return;
        size_t m = flips.size(), m2 = 2 * m;
        double sign = ((m * (m + 1) / 2) % 2 == 0) ? 1 : -1;
        MatC B(ns * m2, 0.), Cinv(m2 * m2, 0.);
        for (size_t k = 0; k < m; k++) Cinv[(m + k) + k * m2] = 1.;
        std::vector<bool> fl(ns, false);
        for (auto f : flips) fl[f] = true;
        for (size_t k = 0; k < m; k++) {
            B[flips[k] + (m + k) * ns] = 1.;
            for (size_t i = 0; i < ns; i++)
                if (flips[k] != i) B[i + k * ns] = a(i, flips[k], s, fl[i], true) - a(i, flips[k], s);
            for (size_t l = 0; l < k; l++)
                Cinv[k + l * m2] = -a(flips[k], flips[l], s, true, true) + a(flips[k], flips[l], s);
        }
        for (size_t i = 0; i < m2; i++)
            for (size_t j = i + 1; j < m2; j++) Cinv[i + j * m2] = -Cinv[j + i * m2];
        MatC invB(ns * m2, 0.);
        for (size_t j = 0; j < m2; j++)
            for (size_t kk = 0; kk < ns; kk++) {
                cd bv = B[kk + j * ns];
                if (bv == 0.) continue;
                for (size_t i = 0; i < ns; i++) invB[i + j * ns] += c.inv[i + kk * ns] * bv;
            }
        MatC tmp(m2 * m2, 0.);
        for (size_t i = 0; i < m2; i++)
            for (size_t j = 0; j < m2; j++) {
                cd acc = 0.;
                for (size_t kk = 0; kk < ns; kk++) acc += B[kk + i * ns] * invB[kk + j * ns];
                tmp[i + j * m2] = 0.5 * acc;
            }
        for (size_t i = 0; i < m2; i++)
            for (size_t j = 0; j < m2; j++) Cinv[i + j * m2] += tmp[i + j * m2] - tmp[j + i * m2];
        MatC Ci;
        if (!inverse(Cinv, m2, Ci)) throw std::runtime_error("singular update");
        MatC C(m2 * m2);
        for (size_t i = 0; i < m2; i++)
            for (size_t j = 0; j < m2; j++) C[i + j * m2] = 0.5 * Ci[i + j * m2];
        MatC Ct = C;
        for (size_t i = 0; i < m2; i++)
            for (size_t j = 0; j < m2; j++) C[i + j * m2] = Ct[i + j * m2] - Ct[j + i * m2];
        MatC t1(ns * m2, 0.);
        for (size_t j = 0; j < m2; j++)
            for (size_t kk = 0; kk < m2; kk++)
                for (size_t i = 0; i < ns; i++) t1[i + j * ns] += invB[i + kk * ns] * C[kk + j * m2];
        MatC t2(ns * ns, 0.);
        for (size_t j = 0; j < ns; j++)
            for (size_t kk = 0; kk < m2; kk++) {
                cd v = invB[j + kk * ns];
                for (size_t i = 0; i < ns; i++) t2[i + j * ns] += 0.5 * t1[i + kk * ns] * v;
            }
        for (size_t i = 0; i < ns; i++)
            for (size_t j = 0; j < ns; j++) c.inv[i + j * ns] += t2[i + j * ns] - t2[j + i * ns];
        int e;
        cd c2 = pfaffian10(Cinv, m2, e) * sign;
        c.pfaff *= c2;
        c.exp += e;
        c.update_factor = c2 * std::pow(10., e);
        int sh = (int)std::log10(std::abs(c.pfaff));
        if (std::abs(sh) > 1) {
            c.pfaff *= std::pow(10., -sh);
            c.exp += sh;
        }
    }
    void derivative(const State& s, const PfCtx& c, cd* out) const {
        for (size_t k = 0; k < n_params(); k++) out[k] = 0.;
        for (size_t i = 0; i < ns; i++)
            for (size_t j = 0; j < i; j++) {
                cd x = c.inv[i + j * ns];
                int bb = b(i, j);
                out[(size_t)(std::abs(bb) - 1) + (size_t)spidx(i, j, s) * rows] += bb > 0 ? -x : x;
            }
    }
};

/* ------------------------------------------------------------- machine --
 * rbm_base.cpp:38-217, 294-368; rbm_symmetry.cpp:31-147; math.cpp:37-80. */

/* sum_i ln cosh(x_i), math.cpp:37-56 (single chunk) */
static cd lncosh_sum(const cd* x, size_t n) {
    double re = 0, im = 0;
    for (size_t i = 0; i < n; i++) {
        double a = std::abs(x[i].real());
        double lr = a - M_LN2;
        double e = std::exp(-2. * a);
        lr += std::log1p(e);
        double s = std::sin(x[i].imag()), c = std::cos(x[i].imag());
        s *= std::copysign((1. - e) / (1. + e), x[i].real());
        lr += 0.5 * std::log(s * s + c * c);
        re += lr;
        im += std::atan2(s, c);
    }
    return {re, im};
}

struct Ctx {
    std::vector<cd> theta; /* n_alpha x n_symm column-major */
    PfCtx pf;
};

struct Machine {
    size_t N, n_vb, n_alpha, n_symm, H, n_rbm;
    std::vector<std::vector<size_t>> perm;
    std::vector<size_t> basis;
    std::vector<cd> v, hb, W; /* W column-major N x n_alpha */
    std::unique_ptr<Pfaffian> pf;

    size_t n_params() const { return n_rbm + (pf ? pf->n_params() : 0); }

    Ctx get_context(const State& s) const { /* rbm_base.cpp:161-172, rbm_symmetry.cpp:48-61 */
        Ctx c;
        c.theta.assign(H, 0.);
        for (size_t sy = 0; sy < n_symm; sy++)
            for (size_t a = 0; a < n_alpha; a++) {
                cd acc = 0.;
                for (size_t f = 0; f < N; f++) acc += W[perm[sy][f] + a * N] * s.sig(f);
                c.theta[a + sy * n_alpha] = acc + hb[a];
            }
        if (pf) c.pf = pf->get_context(s);
        return c;
    }
    void update_context(const State& s, const std::vector<size_t>& flips, Ctx& c) const {
        /* rbm_base.cpp:174-192, rbm_symmetry.cpp:63-95 */
        for (size_t sy = 0; sy < n_symm; sy++)
            for (auto f : flips) {
                double d = s[f] ? -2. : 2.;
                for (size_t a = 0; a < n_alpha; a++) c.theta[a + sy * n_alpha] -= d * W[perm[sy][f] + a * N];
            }
        if (pf) pf->update_context(s, flips, c.pf);
    }
    cd log_bias(const State& s) const { /* rbm_base.cpp:294-296, rbm_symmetry.cpp:138-147 */
        cd r = 0.;
        for (size_t i = 0; i < n_vb; i++)
            for (size_t sy = 0; sy < n_symm; sy++) r += v[i] * s.sig(perm[sy][basis[i]]);
        return r;
    }
    cd log_psi(const State& s, const Ctx& c) const {
        cd r = log_bias(s) + lncosh_sum(c.theta.data(), H);
        if (pf) r += std::log(c.pf.pfaff) + (double)c.pf.exp * std::log(10.);
        return r;
    }
    /* psi(sigma')/psi(sigma); updated receives the context of sigma'
     * (rbm_base.cpp:145-159, 306-344, pfaffian.hpp:73-86) */
    cd psi_over_psi(const State& s, const std::vector<size_t>& flips, const Ctx& c, Ctx& upd,
                    std::unordered_map<State, cd, StateHash>* lut = nullptr) const {
        if (flips.empty()) return 1.;
        cd ret = 0.;
        for (auto f : flips) ret -= (s[f] ? -2. : 2.) * v[f % n_vb];
        State s2 = s;
        s2.flip(flips);
        upd = c;
        update_context(s, flips, upd);
        auto lnc = [&](const Ctx& cc, const State& st) {
            if (lut) {
                auto it = lut->find(st);
                if (it != lut->end()) return it->second;
                cd r = lncosh_sum(cc.theta.data(), H);
                (*lut)[st] = r;
                return r;
            }
            return lncosh_sum(cc.theta.data(), H);
        };
        ret += lnc(upd, s2) - lnc(c, s);
        cd r = std::exp(ret);
        if (pf) r *= upd.pf.update_factor;
        return r;
    }
    void derivative(const State& s, const Ctx& c, cd* out) const {
        /* rbm_base.cpp:194-217, rbm_symmetry.cpp:97-130 */
        for (size_t k = 0; k < n_rbm; k++) out[k] = 0.;
        for (size_t i = 0; i < n_vb; i++)
            for (size_t sy = 0; sy < n_symm; sy++) out[i] += s.sig(perm[sy][basis[i]]);
        std::vector<cd> t(H);
        for (size_t h = 0; h < H; h++) t[h] = std::tanh(c.theta[h]);
        for (size_t a = 0; a < n_alpha; a++)
            for (size_t sy = 0; sy < n_symm; sy++) out[n_vb + a] += t[a + sy * n_alpha];
        std::vector<double> ps(N);
        for (size_t sy = 0; sy < n_symm; sy++) {
            for (size_t f = 0; f < N; f++) ps[perm[sy][f]] = s.sig(f); /* (P_s sigma) */
            for (size_t a = 0; a < n_alpha; a++) {
                cd ta = t[a + sy * n_alpha];
                for (size_t k = 0; k < N; k++) out[n_vb + n_alpha + k + a * N] += ps[k] * ta;
            }
        }
        if (pf) pf->derivative(s, c.pf, out + n_rbm);
    }
    void set_params(const cd* p) {
        std::copy(p, p + n_vb, v.begin());
        std::copy(p + n_vb, p + n_vb + n_alpha, hb.begin());
        std::copy(p + n_vb + n_alpha, p + n_rbm, W.begin());
        if (pf) std::copy(p + n_rbm, p + n_rbm + pf->n_params(), pf->fs.begin());
    }
    void get_params(cd* p) const {
        std::copy(v.begin(), v.end(), p);
        std::copy(hb.begin(), hb.end(), p + n_vb);
        std::copy(W.begin(), W.end(), p + n_vb + n_alpha);
        if (pf) std::copy(pf->fs.begin(), pf->fs.end(), p + n_rbm);
    }
    void update_weights(const cd* dw) { /* rbm_base.cpp:117-143 */
        for (size_t i = 0; i < n_vb; i++) v[i] -= dw[i];
        for (size_t i = 0; i < n_alpha; i++) hb[i] -= dw[n_vb + i];
        for (size_t i = 0; i < N * n_alpha; i++) W[i] -= dw[n_vb + n_alpha + i];
        if (pf)
            for (size_t i = 0; i < pf->n_params(); i++) pf->fs[i] -= dw[n_rbm + i];
    }
};

/* ----------------------------------------------------------- aggregator --
 * aggregator.cpp:29-241 for one rank. */
struct Agg {
    size_t n_samples, dim, n_bins = 0, bin_size = 0, cur_n_bin = 0;
    bool track = false, binning = false;
    double wsum = 0, wsum2 = 0, wsum_bin = 0, sample_factor = 1;
    std::vector<cd> result, result_binned, bin;
    std::vector<double> variance, variance_binned, tau;
    Agg(size_t ns, size_t d) : n_samples(ns), dim(d), result(d, 0.) {}
    void track_variance(size_t nb) {
        if (nb > 1) binning = true;
        track = true;
        n_bins = nb;
        if (binning && n_samples % nb != 0) throw std::runtime_error("n_samples not divisable by n_bins!");
        if (binning) bin_size = n_samples / n_bins;
        bin.assign(dim, 0.);
        result_binned.assign(dim, 0.);
        variance.assign(dim, 0.);
        variance_binned.assign(dim, 0.);
        tau.assign(dim, 0.);
    }
    void aggregate(const cd* x, double w = 1.) {
        wsum += w;
        wsum2 += w * w;
        std::vector<cd> delta(dim);
        for (size_t i = 0; i < dim; i++) {
            delta[i] = w * (x[i] - result[i]);
            result[i] += delta[i] / wsum;
        }
        if (!track) return;
        cur_n_bin++;
        for (size_t i = 0; i < dim; i++) variance[i] += (std::conj(x[i] - result[i]) * delta[i]).real();
        if (binning) {
            wsum_bin += w;
            for (size_t i = 0; i < dim; i++) bin[i] += w * x[i];
            if (cur_n_bin == bin_size) {
                cur_n_bin = 0;
                for (size_t i = 0; i < dim; i++) {
                    bin[i] /= wsum_bin;
                    cd d = wsum_bin * (bin[i] - result_binned[i]);
                    result_binned[i] += d / wsum;
                    variance_binned[i] += (std::conj(d) * (bin[i] - result_binned[i])).real();
                    bin[i] = 0.;
                }
                wsum_bin = 0;
            }
        }
    }
    void finalize(double ptotal) {
        if (track) {
            if (binning) {
                bool zero = true;
                for (size_t i = 0; i < dim; i++) zero &= variance[i] == 0 && variance_binned[i] == 0;
                for (size_t i = 0; i < dim; i++)
                    tau[i] = zero ? 0. : 0.5 * (std::max(bin_size * variance_binned[i] / variance[i], 1.) - 1);
                for (size_t i = 0; i < dim; i++)
                    variance_binned[i] = std::max(variance_binned[i] / (ptotal - wsum2 / ptotal), 0.);
            }
            for (size_t i = 0; i < dim; i++) variance[i] = std::max(variance[i] / (ptotal - wsum2 / ptotal), 0.);
            sample_factor = ptotal / (ptotal - wsum2 / ptotal);
        }
        for (size_t i = 0; i < dim; i++) result[i] *= wsum / ptotal;
    }
    double stddev(size_t i) const {
        return binning ? std::sqrt(variance_binned[i] / n_bins) : std::sqrt(variance[i] / n_samples);
    }
};

static size_t default_bins(size_t n) { /* main.cpp:676-677, abstract_optimizer.cpp:188-189 */
    size_t b = 50;
    while (n % b != 0) b--;
    return b;
}

/* ------------------------------------------------------------------ SR --
 * cg_solver.cpp:119-189, 301-317; SR.cpp:81-116; abstract_optimizer.cpp:27-178 */
struct DecayT {
    double initial, min, decay, value;
    bool min_reached = false;
    DecayT(const Decay& d, size_t offset)
        : initial(d.initial), min(d.min), decay(d.decay), value(std::pow(d.decay, (double)offset) * d.initial) {
        if (decay == 1.) min_reached = true;
    }
    double get() {
        if (!min_reached) {
            value *= decay;
            if (value < min) {
                min_reached = true;
                value = min;
            }
        }
        return value;
    }
};

static void aprod(size_t P, size_t S, size_t n_neural, const cd* O, const cd* d, const double* diag,
                  double r1, double r2, double rd, const cd* b, cd* x) {
    double maxd = 0;
    for (size_t p = 0; p < P; p++) maxd = std::max(maxd, diag[p]);
    std::vector<cd> y(S, 0.);
    /* every y_k and x_p is summed by one thread in a fixed order, so the
     * result does not depend on the thread count */
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < S; k++) {
        cd acc = 0.;
        for (size_t p = 0; p < P; p++) acc += O[p + k * P] * b[p];
        y[k] = acc;
    }
    cd db = 0.;
    for (size_t p = 0; p < P; p++) db += d[p] * b[p];
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < P; p++) {
        double reg = diag[p] * (p < n_neural ? r1 : r1 + rd) + maxd * r2;
        cd acc = 0.;
        for (size_t k = 0; k < S; k++) acc += std::conj(O[p + k * P]) * y[k];
        x[p] = reg * b[p] + acc / (double)S - std::conj(d[p]) * db;
    }
}

static void cg1(size_t P, size_t S, size_t n_neural, const cd* O, const cd* d, const double* diag,
                double r1, double r2, double rd, const cd* b, cd* x, size_t max_iter, double rtol,
                size_t& itn, double& rs) {
    std::vector<cd> r(b, b + P), Ap(P), p(P);
    aprod(P, S, n_neural, O, d, diag, r1, r2, rd, x, Ap.data());
    for (size_t i = 0; i < P; i++) r[i] -= Ap[i];
    p = r;
    double rsold = 0;
    for (auto& v : r) rsold += std::norm(v);
    rs = rsold;
    for (itn = 0; itn < max_iter; itn++) {
        aprod(P, S, n_neural, O, d, diag, r1, r2, rd, p.data(), Ap.data());
        cd alpha = 0.;
        for (size_t i = 0; i < P; i++) alpha += std::conj(p[i]) * Ap[i];
        alpha = rsold / alpha;
        for (size_t i = 0; i < P; i++) x[i] += alpha * p[i];
        for (size_t i = 0; i < P; i++) r[i] -= alpha * Ap[i];
        rs = 0;
        for (auto& v : r) rs += std::norm(v);
        if (rs < rtol) break;
        for (size_t i = 0; i < P; i++) p[i] = r[i] + (rs / rsold) * p[i];
        rsold = rs;
    }
}

/* ------------------------------------------------------------ container */
struct Sys {
    Model model;
    Machine mach;
    std::vector<size_t> pba, pbb; /* proposal bonds (type != 2) */
    size_t nw;
    /* optimiser state (abstract_optimizer / SR) */
    std::unique_ptr<DecayT> lr, k1, k2, kd;
    std::vector<cd> dw;
    size_t n_updates = 0;
};

static void build_sys(Sys& S, const std::string& ini) {
    S.model.cfg = parse_ini(ini);
    build_model(S.model);
    const Config& c = S.model.cfg;
    Lattice& lat = S.model.lat;
    Machine& m = S.mach;
    m.N = lat.n_total;
    if (m.N > 128) throw std::runtime_error("n must be smaller than MAX_SPIN_SITES!");
    if (c.rbm == "basic") {
        m.n_vb = m.N;
        m.n_symm = 1;
        m.perm = lat.symmetry({0});
        m.basis.resize(m.N);
        for (size_t i = 0; i < m.N; i++) m.basis[i] = i;
    } else if (c.rbm == "symmetry") {
        const auto& sym = lat.default_symm;
        m.n_symm = lat.symmetry_size(sym);
        m.n_vb = m.N / m.n_symm;
        m.perm = lat.symmetry(sym);
        m.basis = lat.symm_basis(sym);
        if (m.perm.size() != m.n_symm) throw std::runtime_error("symmetry size mismatch");
    } else {
        throw std::runtime_error("RBM Type '" + c.rbm + "' not available!");
    }
    m.n_alpha = c.alpha * m.n_vb;
    m.H = m.n_alpha * m.n_symm;
    m.n_rbm = m.n_vb + (c.alpha + c.alpha * m.N) * m.n_vb;
    m.v.assign(m.n_vb, 0.);
    m.hb.assign(m.n_alpha, 0.);
    m.W.assign(m.N * m.n_alpha, 0.);
    if (c.pfaffian) {
        m.pf.reset(new Pfaffian());
        m.pf->build(lat, c.pf_symmetry);
    }
    for (auto& b : lat.bonds)
        if (b[2] != 2) S.pba.push_back(b[0]), S.pbb.push_back(b[1]);
    S.nw = (m.N + 63) / 64;
    S.lr.reset(new DecayT(c.lr, 0));
    S.k1.reset(new DecayT(c.reg1, 0));
    S.k2.reset(new DecayT(c.reg2, 0));
    S.kd.reset(new DecayT(c.dreg1, 0));
    S.dw.assign(m.n_params(), 0.);
}

/* local energy, local_op.cpp:31-60 summed as local_op_chain.cpp:54-60 */
static cd eloc(const Sys& S, const State& s, const Ctx& c,
               std::unordered_map<State, cd, StateHash>* lut = nullptr) {
    cd e = 0.;
    Ctx tmp;
    std::vector<size_t> flips;
    for (auto& op : S.model.ham) {
        size_t loc = s.to_num_loc(op.sites);
        for (size_t j = 0; j < op.dim; j++) {
            cd val = op.mat[loc * op.dim + j];
            if (val == 0.) continue;
            if (j == loc) {
                e += val;
            } else {
                flips.clear();
                size_t x = j ^ loc;
                for (size_t i = 0; i < op.sites.size(); i++)
                    if ((x >> i) & 1) flips.push_back(op.sites[i]);
                e += val * S.mach.psi_over_psi(s, flips, c, tmp, lut);
            }
        }
    }
    return e;
}

/* one Markov chain, metropolis_sampler.cpp:118-273 */
static double sample_chain(const Sys& S, uint64_t seed, uint32_t epoch, size_t chain, size_t ns,
                           bool use_lut, uint64_t* st_out, cd* e_out, cd* o_out) {
    const Config& cfg = S.model.cfg;
    const Machine& m = S.mach;
    size_t total = ns * cfg.step + cfg.warmup;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)chain, epoch, 0xFFFFFFFFu, 0}, r[4];
    philox(ctr, key, r);
    State s;
    for (size_t i = 0; i < m.N; i++)
        if ((r[i >> 5] >> (i & 31)) & 1) s.flip(i);
    std::unordered_map<State, cd, StateHash> lut;
    Ctx c = m.get_context(s), nc;
    uint64_t thr = (uint64_t)std::min(std::floor(cfg.bond_flips * 4294967296.0), 4294967296.0);
    size_t acc = 0, proposed = 0, k = 0;
    std::vector<size_t> flips;
    for (size_t step = 0; step < total; step++) {
        for (size_t sw = 0; sw < m.N; sw++) {
            uint32_t cc[4] = {(uint32_t)chain, epoch, (uint32_t)step, (uint32_t)sw};
            philox(cc, key, r);
            flips.clear();
            if ((uint64_t)r[0] < thr && !S.pba.empty()) {
                size_t bi = (size_t)(((uint64_t)r[1] * S.pba.size()) >> 32);
                flips = {S.pba[bi], S.pbb[bi]};
                sw++;
            } else {
                flips.push_back((size_t)(((uint64_t)r[1] * m.N) >> 32));
            }
            proposed++;
            cd ratio = m.psi_over_psi(s, flips, c, nc, use_lut ? &lut : nullptr);
            double a = std::norm(ratio);
            if (u53(r[2], r[3]) < a) {
                std::swap(c, nc);
                acc++;
                if (m.pf && cfg.pf_refresh && acc % (size_t)cfg.pf_refresh == 0) {
                    State s2 = s;
                    s2.flip(flips);
                    c.pf = m.pf->get_context(s2); /* abstract_sampler.cpp:103-113 */
                }
                s.flip(flips);
            }
        }
        if (step >= cfg.warmup && (step - cfg.warmup) % cfg.step == 0) {
            if (st_out)
                for (size_t w = 0; w < S.nw; w++) st_out[k * S.nw + w] = s.w[w];
            if (e_out) e_out[k] = eloc(S, s, c, use_lut ? &lut : nullptr);
            if (o_out) m.derivative(s, c, o_out + k * m.n_params());
            k++;
        }
    }
    return (double)acc / (double)proposed;
}

static void sample_all(Sys& S, uint64_t seed, uint32_t epoch, int threads, bool use_lut, uint64_t* st,
                       cd* e, cd* O, double* cacc) {
    const Config& cfg = S.model.cfg;
    size_t spc = cfg.n_samples / cfg.n_chains, res = cfg.n_samples - spc * cfg.n_chains;
    std::vector<size_t> off(cfg.n_chains + 1, 0);
    for (size_t c = 0; c < cfg.n_chains; c++) off[c + 1] = off[c] + spc + (c == 0 ? res : 0);
    size_t P = S.mach.n_params();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long c = 0; c < (long)cfg.n_chains; c++) {
        double a = sample_chain(S, seed, epoch, (size_t)c, off[c + 1] - off[c], use_lut,
                                st ? st + off[c] * S.nw : nullptr, e ? e + off[c] : nullptr,
                                O ? O + off[c] * P : nullptr);
        if (cacc) cacc[c] = a;
    }
}

}  // namespace orc

using namespace orc;

struct orc_model {
    Sys s;
};

static int fail(const std::exception& e) {
    g_err = e.what();
    return -1;
}
#define GUARD(...)                       \
    try {                                \
        __VA_ARGS__;                     \
    } catch (const std::exception& e) { \
        return fail(e);                  \
    }                                    \
    return 0;

extern "C" {

const char* orc_error(void) { return g_err.c_str(); }

orc_model* orc_create(const char* ini) {
    try {
        auto* m = new orc_model();
        build_sys(m->s, ini);
        return m;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_destroy(orc_model* m) { delete m; }

int orc_info(const orc_model* m, int64_t* o) {
    const Machine& mc = m->s.mach;
    o[0] = mc.N;
    o[1] = mc.n_params();
    o[2] = mc.H;
    o[3] = mc.n_symm;
    o[4] = mc.n_vb;
    o[5] = mc.n_alpha;
    o[6] = mc.pf ? mc.pf->n_params() : 0;
    o[7] = m->s.model.lat.bonds.size();
    o[8] = m->s.model.ham.size();
    o[9] = m->s.nw;
    o[10] = mc.n_rbm;
    return 0;
}
int orc_bonds(const orc_model* m, int64_t* out) {
    size_t i = 0;
    for (auto& b : m->s.model.lat.bonds) out[i++] = b[0], out[i++] = b[1], out[i++] = b[2];
    return 0;
}
int orc_nns(const orc_model* m, int64_t site, int64_t* out) {
    auto nn = m->s.model.lat.nns((size_t)site);
    for (size_t i = 0; i < nn.size(); i++) out[i] = nn[i];
    return 0;
}
int orc_hexagons(const orc_model* m, int64_t* out, int64_t* n) {
    if (m->s.model.lat.kind != 0) return -1;
    auto hx = m->s.model.lat.hexagons();
    size_t i = 0;
    for (auto& h : hx)
        for (auto x : h) out[i++] = x;
    *n = hx.size();
    return 0;
}
int orc_plaqs(const orc_model* m, int64_t* out, int64_t* n) {
    if (m->s.model.lat.kind != 1) return -1;
    auto pl = m->s.model.lat.plaqs();
    size_t i = 0;
    for (auto& p : pl) {
        for (auto x : p.first) out[i++] = x;
        out[i++] = p.second;
    }
    *n = pl.size();
    return 0;
}
int orc_pf_map(const orc_model* m, int64_t* out, int64_t* rows) {
    const Pfaffian* pf = m->s.mach.pf.get();
    if (!pf) return -1;
    for (size_t i = 0; i < pf->bs.size(); i++) out[i] = pf->bs[i];
    *rows = (int64_t)pf->rows;
    return 0;
}
int orc_symmetry(const orc_model* m, int64_t* perms) {
    const Machine& mc = m->s.mach;
    for (size_t s = 0; s < mc.n_symm; s++)
        for (size_t f = 0; f < mc.N; f++) perms[s * mc.N + f] = mc.perm[s][f];
    return 0;
}
int orc_up_down(const orc_model* m, int64_t uc, int64_t dir, int64_t step, int64_t* up, int64_t* down) {
    *up = m->s.model.lat.up(uc, dir, step);
    *down = m->s.model.lat.down(uc, dir, step);
    return 0;
}
int orc_bravais_updown(int64_t n_uc, int64_t n_uc_b, int64_t h_shift, int64_t uc, int64_t dir,
                       int64_t step, int64_t* up, int64_t* down) {
    Lattice l; /* the bravais_test stub of test_lattice.cpp:19-51 */
    l.n_uc = n_uc;
    l.n_uc_b = n_uc_b == 0 ? n_uc : n_uc_b;
    l.h_shift = h_shift;
    l.n_total_uc = n_uc * l.n_uc_b;
    *up = l.up(uc, dir, step);
    *down = l.down(uc, dir, step);
    return 0;
}
double orc_exact_energy(const orc_model* m) { return m->s.model.exact_energy(); }

int orc_term_info(const orc_model* m, int64_t t, int64_t* k) {
    *k = m->s.model.ham.at(t).sites.size();
    return 0;
}
int orc_term(const orc_model* m, int64_t t, int64_t* sites, double* mat) {
    auto& op = m->s.model.ham.at(t);
    for (size_t i = 0; i < op.sites.size(); i++) sites[i] = op.sites[i];
    std::memcpy(mat, op.mat.data(), op.mat.size() * sizeof(cd));
    return 0;
}

int orc_set_params(orc_model* m, const double* p) {
    m->s.mach.set_params(reinterpret_cast<const cd*>(p));
    return 0;
}
int orc_get_params(const orc_model* m, double* p) {
    m->s.mach.get_params(reinterpret_cast<cd*>(p));
    return 0;
}
int orc_theta(const orc_model* m, const uint64_t* st, double* th) {
    GUARD({
        auto c = m->s.mach.get_context(load_state(st, m->s.nw));
        std::memcpy(th, c.theta.data(), c.theta.size() * sizeof(cd));
    })
}
int orc_logpsi(const orc_model* m, const uint64_t* st, int64_t n, double* out) {
    GUARD({
        cd* o = reinterpret_cast<cd*>(out);
        for (int64_t i = 0; i < n; i++) {
            State s = load_state(st + i * m->s.nw, m->s.nw);
            o[i] = m->s.mach.log_psi(s, m->s.mach.get_context(s));
        }
    })
}
int orc_psi_ratio(const orc_model* m, const uint64_t* st, const int64_t* flips, int64_t nf, double* out) {
    GUARD({
        State s = load_state(st, m->s.nw);
        std::vector<size_t> f(flips, flips + nf);
        Ctx c = m->s.mach.get_context(s), u;
        *reinterpret_cast<cd*>(out) = m->s.mach.psi_over_psi(s, f, c, u);
    })
}
int orc_eloc(const orc_model* m, const uint64_t* st, int64_t n, double* out) {
    GUARD({
        cd* o = reinterpret_cast<cd*>(out);
        for (int64_t i = 0; i < n; i++) {
            State s = load_state(st + i * m->s.nw, m->s.nw);
            o[i] = eloc(m->s, s, m->s.mach.get_context(s));
        }
    })
}
int orc_derivative(const orc_model* m, const uint64_t* st, int64_t n, double* out) {
    GUARD({
        cd* o = reinterpret_cast<cd*>(out);
        size_t P = m->s.mach.n_params();
        for (int64_t i = 0; i < n; i++) {
            State s = load_state(st + i * m->s.nw, m->s.nw);
            m->s.mach.derivative(s, m->s.mach.get_context(s), o + i * P);
        }
    })
}
int orc_update_check(const orc_model* m, const uint64_t* st, const int64_t* flips, int64_t n_steps,
                     int64_t per, double* diff) {
    GUARD({
        const Machine& mc = m->s.mach;
        State s = load_state(st, m->s.nw);
        Ctx c = mc.get_context(s);
        for (int64_t i = 0; i < n_steps; i++) {
            std::vector<size_t> f(flips + i * per, flips + (i + 1) * per);
            mc.update_context(s, f, c);
            s.flip(f);
        }
        Ctx c2 = mc.get_context(s);
        double d = 0;
        for (size_t h = 0; h < mc.H; h++) d += std::abs(c.theta[h] - c2.theta[h]);
        *diff = d / mc.H;
    })
}

int orc_pfaffian_dense(const double* A, int64_t n, double* mant, int64_t* e10) {
    GUARD({
        MatC a(reinterpret_cast<const cd*>(A), reinterpret_cast<const cd*>(A) + n * n);
        int e;
        cd p = pfaffian10(a, (size_t)n, e);
        mant[0] = p.real();
        mant[1] = p.imag();
        *e10 = e;
    })
}
int orc_pf_context(const orc_model* m, const uint64_t* st, double* inv, double* mant, int64_t* e10) {
    GUARD({
        if (!m->s.mach.pf) throw std::runtime_error("no Pfaffian");
        PfCtx c = m->s.mach.pf->get_context(load_state(st, m->s.nw));
        if (inv) std::memcpy(inv, c.inv.data(), c.inv.size() * sizeof(cd));
        mant[0] = c.pfaff.real();
        mant[1] = c.pfaff.imag();
        *e10 = c.exp;
    })
}
int orc_pf_update_check(const orc_model* m, const uint64_t* st, const int64_t* flips, int64_t n_steps,
                        int64_t per, double* inv_diff, double* pf_diff, double* ratio_out) {
    GUARD({
        const Pfaffian* pf = m->s.mach.pf.get();
        if (!pf) throw std::runtime_error("no Pfaffian");
        State s = load_state(st, m->s.nw);
        PfCtx c = pf->get_context(s);
        for (int64_t i = 0; i < n_steps; i++) {
            std::vector<size_t> f(flips + i * per, flips + (i + 1) * per);
            pf->update_context(s, f, c);
            if (ratio_out) {
                ratio_out[2 * i] = c.update_factor.real();
                ratio_out[2 * i + 1] = c.update_factor.imag();
            }
            s.flip(f);
        }
        PfCtx c2 = pf->get_context(s);
        double d = 0;
        for (size_t k = 0; k < c.inv.size(); k++) d += std::abs(c.inv[k] - c2.inv[k]);
        *inv_diff = d / c.inv.size();
        *pf_diff = std::abs(c.pfaff - c2.pfaff * std::pow(10., c2.exp - c.exp)) * std::pow(10., c.exp);
    })
}

int orc_sample(orc_model* m, uint64_t seed, uint32_t epoch, int threads, int use_lut, uint64_t* st,
               double* e, double* O, double* cacc) {
    GUARD(sample_all(m->s, seed, epoch, threads, use_lut != 0, st, reinterpret_cast<cd*>(e),
                     reinterpret_cast<cd*>(O), cacc))
}

int orc_aggregate(const orc_model* m, const double* e_, const double* O_, int64_t S, int64_t n_bins,
                  double* oe, double* oo) {
    GUARD({
        const cd* e = reinterpret_cast<const cd*>(e_);
        const cd* O = reinterpret_cast<const cd*>(O_);
        size_t P = m->s.mach.n_params();
        size_t nb = n_bins > 0 ? (size_t)n_bins : default_bins((size_t)S);
        Agg ah(S, 1);
        ah.track_variance(nb);
        for (int64_t k = 0; k < S; k++) ah.aggregate(e + k);
        ah.finalize((double)S);
        oe[0] = ah.result[0].real();
        oe[1] = ah.result[0].imag();
        oe[2] = ah.variance[0];
        oe[3] = ah.variance_binned[0];
        oe[4] = ah.tau[0];
        oe[5] = ah.stddev(0);
        if (O && oo) {
            Agg ad(S, P), adh(S, P);
            ad.track_variance(0);
            std::vector<cd> x(P);
            for (int64_t k = 0; k < S; k++) {
                ad.aggregate(O + k * P);
                for (size_t p = 0; p < P; p++) x[p] = e[k] * std::conj(O[k * P + p]);
                adh.aggregate(x.data());
            }
            ad.finalize((double)S);
            adh.finalize((double)S);
            for (size_t p = 0; p < P; p++) {
                oo[2 * p] = ad.result[p].real();
                oo[2 * p + 1] = ad.result[p].imag();
                oo[2 * P + p] = ad.variance[p];
                oo[3 * P + 2 * p] = adh.result[p].real();
                oo[3 * P + 2 * p + 1] = adh.result[p].imag();
            }
        }
    })
}

int orc_apply_S(int64_t P, int64_t S, int64_t nn, const double* O, const double* mean, const double* var,
                double r1, double r2, double rd, const double* x, double* y) {
    aprod(P, S, nn, reinterpret_cast<const cd*>(O), reinterpret_cast<const cd*>(mean), var, r1, r2, rd,
          reinterpret_cast<const cd*>(x), reinterpret_cast<cd*>(y));
    return 0;
}
int orc_cg(int64_t P, int64_t S, int64_t nn, const double* O, const double* mean, const double* var,
           double r1, double r2, double rd, const double* b, double* x, int64_t max_iter, double rtol,
           int64_t* itn, double* rs) {
    size_t it;
    cg1(P, S, nn, reinterpret_cast<const cd*>(O), reinterpret_cast<const cd*>(mean), var, r1, r2, rd,
        reinterpret_cast<const cd*>(b), reinterpret_cast<cd*>(x), max_iter, rtol, it, *rs);
    *itn = it;
    return 0;
}

int orc_epoch(orc_model* m, uint64_t seed, uint32_t epoch, int threads, int use_lut, double* stats) {
    GUARD({
        Sys& s = m->s;
        const Config& c = s.model.cfg;
        size_t S = c.n_samples, P = s.mach.n_params();
        std::vector<cd> e(S), O(P * S);
        auto t0 = std::chrono::steady_clock::now();
        std::vector<double> cacc(c.n_chains);
        sample_all(s, seed, epoch, threads, use_lut != 0, nullptr, e.data(), O.data(), cacc.data());
        double acc = 0;
        for (double a : cacc) acc += a;
        acc /= c.n_chains;
        auto t1 = std::chrono::steady_clock::now();
        /* aggregators registered by abstract_optimizer.cpp:184-192, SR.cpp:74-79 */
        Agg ah(S, 1), ad(S, P), adh(S, P);
        ah.track_variance(default_bins(S));
        ad.track_variance(0);
        std::vector<cd> x(P);
        for (size_t k = 0; k < S; k++) {
            ah.aggregate(&e[k]);
            ad.aggregate(&O[k * P]);
            for (size_t p = 0; p < P; p++) x[p] = e[k] * std::conj(O[k * P + p]);
            adh.aggregate(x.data());
        }
        ah.finalize(S);
        ad.finalize(S);
        adh.finalize(S);
        double lr = s.lr->get();
        double r1 = s.k1->get(), r2 = s.k2->get(), rd = s.kd->get();
        std::vector<cd> F(P);
        for (size_t p = 0; p < P; p++) F[p] = adh.result[p] - ah.result[0] * std::conj(ad.result[p]);
        size_t mi = std::min(2 * P, c.sr_max_iter);
        if (mi == 0) mi = 2 * P;
        double rtol = c.sr_rtol == 0. ? 1e-16 : c.sr_rtol;
        size_t itn;
        double rs;
        if (threads > 0) omp_set_num_threads(threads);  // the S.x loops of the CG
        cg1(P, S, s.mach.n_rbm, O.data(), ad.result.data(), ad.variance.data(), r1, r2, rd, F.data(),
            s.dw.data(), mi, rtol, itn, rs);
        double mx = 0;
        for (auto& v : s.dw) mx = std::max(mx, std::abs(v));
        bool ok = mx <= 30;
        if (ok) {
            for (auto& v : s.dw) v *= lr;
            s.mach.update_weights(s.dw.data());
            s.n_updates++;
        }
        auto t2 = std::chrono::steady_clock::now();
        size_t N = s.mach.N;
        stats[0] = ah.result[0].real() / N;
        stats[1] = ah.stddev(0) / N;
        stats[2] = acc;
        stats[3] = ah.tau[0];
        stats[4] = (double)itn;
        stats[5] = rs;
        stats[6] = mx;
        stats[7] = ok ? 1 : 0;
        stats[8] = std::chrono::duration<double>(t1 - t0).count();
        stats[9] = std::chrono::duration<double>(t2 - t1).count();
    })
}

int orc_exact_expectation(const orc_model* m, double* out) {
    GUARD({
        const Machine& mc = m->s.mach;
        if (mc.N > 24) throw std::runtime_error("N too large for enumeration");
        size_t n = (size_t)1 << mc.N;
        std::vector<double> lp(n);
        std::vector<cd> el(n);
        double mx = -1e300;
        for (size_t i = 0; i < n; i++) {
            State s;
            s.w[0] = i;
            Ctx c = mc.get_context(s);
            lp[i] = 2 * mc.log_psi(s, c).real();
            mx = std::max(mx, lp[i]);
            el[i] = eloc(m->s, s, c);
        }
        double z = 0;
        cd e = 0.;
        for (size_t i = 0; i < n; i++) {
            double w = std::exp(lp[i] - mx);
            z += w;
            e += w * el[i];
        }
        e /= z;
        out[0] = e.real();
        out[1] = e.imag();
    })
}

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

/* ---------------------------------------------------- learned proposal --
 * No counterpart in the reference (its proposals are single and bond flips,
 * metropolis_sampler.cpp:195-219).  This restates the build's own design:
 *
 * Cluster catalogue: the loops of the lattice (Kitaev: hexagons,
 * honeycomb.cpp:201-209; toric: vertex stars, type 1 of
 * toric_lattice.cpp:57-69).  For site i, cyc(i) is the first loop in list
 * order that contains i, rotated to start at i; cluster (i, k) is its first
 * k sites, k = 1 .. loop length. */
int orc_clusters(const orc_model* m, int64_t* cyc, int64_t* kmax) {
    GUARD({
        const Model& md = m->s.model;
        std::vector<std::vector<size_t>> loops;
        if (md.cfg.model == "kitaev") {
            loops = md.lat.hexagons();
        } else {
            for (auto& p : md.lat.plaqs())
                if (p.second == 1) loops.push_back({p.first[0], p.first[1], p.first[2], p.first[3]});
        }
        size_t K = loops.at(0).size(), N = md.lat.n_total;
        *kmax = (int64_t)K;
        for (size_t i = 0; i < N; i++) {
            bool found = false;
            for (auto& l : loops) {
                auto it = std::find(l.begin(), l.end(), i);
                if (it == l.end()) continue;
                size_t p0 = (size_t)(it - l.begin());
                for (size_t t = 0; t < K; t++) cyc[i * K + t] = (int64_t)l[(p0 + t) % K];
                found = true;
                break;
            }
            if (!found) throw std::runtime_error("site on no loop");
        }
    })
}

/* Proposal network z(s) = W2 relu(W1 s + b1) + b2 with s_j = +-1 (bit set
 * <=> -1), W1 row-major Hn x N, W2 row-major N x Hn; FP64. */
static void mlp_fwd(int64_t N, int64_t Hn, const double* W1, const double* b1, const double* W2, const double* b2,
                    const double* s, double* u, double* z) {
    for (int64_t h = 0; h < Hn; h++) {
        double acc = b1[h];
        for (int64_t j = 0; j < N; j++) acc += W1[h * N + j] * s[j];
        u[h] = acc;
    }
    for (int64_t o = 0; o < N; o++) {
        double acc = b2[o];
        for (int64_t h = 0; h < Hn; h++) acc += W2[o * Hn + h] * std::max(u[h], 0.);
        z[o] = acc;
    }
}

int orc_mlp_logits(int64_t N, int64_t Hn, const double* W1, const double* b1, const double* W2, const double* b2,
                   const uint64_t* st, int64_t nw, int64_t n, double* z) {
    GUARD({
        std::vector<double> s(N), u(Hn);
        for (int64_t r = 0; r < n; r++) {
            for (int64_t j = 0; j < N; j++) s[j] = ((st[r * nw + (j >> 6)] >> (j & 63)) & 1) ? -1. : 1.;
            mlp_fwd(N, Hn, W1, b1, W2, b2, s.data(), u.data(), z + r * N);
        }
    })
}

/* Batch loss L = (1/n) sum_r [ -w_r ln p_{i_r}(s_r) - beta H(p(s_r)) ],
 * p = softmax(z), H = -sum p ln p, and its gradient by backpropagation:
 * dL/dz_j = (1/n) [ w_r (p_j - [j = i_r]) + beta p_j (ln p_j + H) ]. */
int orc_mlp_grad(int64_t N, int64_t Hn, const double* W1, const double* b1, const double* W2, const double* b2,
                 const uint64_t* st, int64_t nw, int64_t n, const int64_t* seed, const double* w, double beta,
                 double* gW1, double* gb1, double* gW2, double* gb2, double* loss) {
    GUARD({
        std::fill(gW1, gW1 + Hn * N, 0.);
        std::fill(gb1, gb1 + Hn, 0.);
        std::fill(gW2, gW2 + N * Hn, 0.);
        std::fill(gb2, gb2 + N, 0.);
        std::vector<double> s(N), u(Hn), z(N), p(N), dz(N), da(Hn);
        double L = 0;
        for (int64_t r = 0; r < n; r++) {
            for (int64_t j = 0; j < N; j++) s[j] = ((st[r * nw + (j >> 6)] >> (j & 63)) & 1) ? -1. : 1.;
            mlp_fwd(N, Hn, W1, b1, W2, b2, s.data(), u.data(), z.data());
            double mx = *std::max_element(z.begin(), z.end()), zs = 0;
            for (int64_t o = 0; o < N; o++) zs += std::exp(z[o] - mx);
            double lse = mx + std::log(zs), H = 0;
            for (int64_t o = 0; o < N; o++) {
                p[o] = std::exp(z[o] - lse);
                H -= p[o] * (z[o] - lse);
            }
            L += -w[r] * (z[seed[r]] - lse) - beta * H;
            for (int64_t o = 0; o < N; o++)
                dz[o] = (w[r] * (p[o] - (o == seed[r] ? 1. : 0.)) + beta * p[o] * ((z[o] - lse) + H)) / (double)n;
            for (int64_t h = 0; h < Hn; h++) {
                double a = std::max(u[h], 0.), acc = 0;
                for (int64_t o = 0; o < N; o++) {
                    gW2[o * Hn + h] += dz[o] * a;
                    acc += W2[o * Hn + h] * dz[o];
                }
                da[h] = u[h] > 0 ? acc : 0.;
            }
            for (int64_t o = 0; o < N; o++) gb2[o] += dz[o];
            for (int64_t h = 0; h < Hn; h++) {
                gb1[h] += da[h];
                for (int64_t j = 0; j < N; j++) gW1[h * N + j] += da[h] * s[j];
            }
        }
        *loss = L / (double)n;
    })
}

}  // extern "C"
