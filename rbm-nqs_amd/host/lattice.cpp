// lattice.cpp -- see lattice.hpp.  Index arithmetic follows
// src/lattice/bravais.cpp, honeycomb.cpp and toric_lattice.cpp exactly, since
// bond lists, hexagons and permutations are part of the model definition.
#include "lattice.hpp"

#include <cmath>
#include <functional>
#include <stdexcept>

namespace lattice {

static size_t ipow(size_t b, size_t e) {
    size_t r = 1;
    while (e--) r *= b;
    return r;
}

bravais::bravais(size_t n_uc_, size_t n_dim_, size_t n_basis_, size_t n_coord_, size_t n_uc_b_,
                 size_t h_shift_, std::vector<double> default_symmetry)
    : n_uc(n_uc_),
      n_dim(n_dim_),
      n_basis(n_basis_),
      n_coordination(n_coord_),
      n_total_uc(n_uc_b_ == 0 ? ipow(n_uc_, n_dim_) : n_uc_ * n_uc_b_),
      n_total((n_uc_b_ == 0 ? ipow(n_uc_, n_dim_) : n_uc_ * n_uc_b_) * n_basis_),
      n_uc_b(n_uc_b_ == 0 ? n_uc_ : n_uc_b_),
      h_shift(h_shift_),
      default_symmetry_(std::move(default_symmetry)) {}

// The cell index is row-major with the first direction (extent n_uc_b)
// fastest; a shift wraps inside its row and, with h_shift, a wrap in
// direction 1 shifts horizontally (bravais.cpp:66-105).
size_t bravais::up(size_t uc, size_t dir, size_t step) const {
    if (step == 0) return uc;
    const size_t span = ipow(n_uc, dir) * n_uc_b;  // cells in one period along dir
    const size_t ext = dir == 0 ? n_uc_b : n_uc;
    const size_t base = (uc / span) * span;
    const size_t wrapped = (uc + (step * span) / ext) / span;
    step %= ext;
    size_t r = base + (uc + step * span / ext) % span;
    if (wrapped && h_shift && dir == 1) r = down(r, 0, h_shift);
    return r;
}

size_t bravais::down(size_t uc, size_t dir, size_t step) const {
    if (step == 0) return uc;
    const size_t span = ipow(n_uc, dir) * n_uc_b;
    const size_t ext = dir == 0 ? n_uc_b : n_uc;
    const size_t base = (uc / span) * span;
    const size_t not_wrapped = (uc + (ext - step) * span / ext) / span;
    step %= ext;
    size_t r = base + (uc + (ext - step) * span / ext) % span;
    if (!not_wrapped && h_shift && dir == 1) r = up(r, 0, h_shift);
    return r;
}

void bravais::build_bonds() {
    bonds_.clear();
    for (size_t c = 0; c < n_total_uc; c++) {
        const size_t a = idx(c, 0);
        const auto nb = nns(a);
        for (size_t t = 0; t < n_coordination; t++) bonds_.push_back({a, nb[t], t});
    }
}

// symmetry factors -> translation steps per direction (bravais.cpp:249-284)
std::vector<size_t> bravais::clean_symms(const std::vector<double>& symm) const {
    int factor = 1;
    for (double s : symm) {
        if (s < 0) throw std::runtime_error("Negative Symmetry factors are not allowed!");
        if (std::abs(factor * s - int(factor * s)) > 1e-14) {
            double frac = factor * s - std::floor(factor * s);
            factor *= (int)(1 / frac);
        }
    }
    if (symm.size() != 1 && symm.size() != n_dim)
        throw std::runtime_error("Symmetry factor sizes incompatible with lattice dimensions.");
    std::vector<size_t> x(n_dim);
    for (size_t i = 0; i < symm.size(); i++) {
        x[i] = (size_t)int(factor * symm[i]);
        if (x[i] == 0) x[i] = i == 0 ? n_uc_b : n_uc;
        const size_t ext = i == 0 ? n_uc_b : n_uc;
        if (ext % x[i] != 0)
            throw std::runtime_error("Symmetry factor " + std::to_string(x[i]) + " in direction " +
                                     std::to_string(i) + " invalid!");
    }
    if (symm.size() == 1)
        for (size_t i = 1; i < n_dim; i++) x[i] = x[0];
    return x;
}

// all translations by multiples of the steps; direction 0 varies fastest
std::vector<std::vector<size_t>> bravais::construct_uc_symmetry(const std::vector<double>& symm) const {
    std::vector<std::vector<size_t>> out;
    if (no_symmetry(symm)) {
        std::vector<size_t> id(n_total_uc);
        for (size_t i = 0; i < n_total_uc; i++) id[i] = i;
        out.push_back(id);
        return out;
    }
    const auto steps = clean_symms(symm);
    std::vector<size_t> shift(n_dim, 0);
    std::function<void(size_t)> rec = [&](size_t d) {
        const size_t ext = d == 0 ? n_uc_b : n_uc;
        for (size_t i = 0; i < ext; i += steps[d]) {
            shift[d] = i;
            if (d == 0) {
                std::vector<size_t> map(n_total_uc);
                for (size_t c = 0; c < n_total_uc; c++) {
                    size_t t = c;
                    for (size_t k = 0; k < n_dim; k++) t = up(t, k, shift[k]);
                    map[c] = t;
                }
                out.push_back(map);
            } else {
                rec(d - 1);
            }
        }
    };
    rec(n_dim - 1);
    return out;
}

std::vector<perm_t> bravais::construct_symmetry(const std::vector<double>& symm) const {
    std::vector<perm_t> out;
    for (const auto& map : construct_uc_symmetry(symm)) {
        perm_t p(n_total);
        for (size_t i = 0; i < n_total; i++) p[i] = idx(map[uc_idx(i)], b_idx(i));
        out.push_back(p);
    }
    return out;
}

std::vector<size_t> bravais::construct_symm_basis(const std::vector<double>& symm) const {
    std::vector<size_t> out;
    if (no_symmetry(symm)) {
        for (size_t i = 0; i < n_total; i++) out.push_back(i);
        return out;
    }
    // cells of the translation-reduced unit cell (bravais.cpp:207-235)
    const auto steps = clean_symms(symm);
    std::vector<size_t> cells;
    std::function<void(size_t, size_t)> rec = [&](size_t d, size_t c) {
        for (size_t i = 0; i < steps[d]; i++) {
            size_t t = up(c, d, i);
            if (d == 0) cells.push_back(t);
            else rec(d - 1, t);
        }
    };
    rec(n_dim - 1, 0);
    for (size_t c : cells)
        for (size_t b = 0; b < n_basis; b++) out.push_back(idx(c, b));
    return out;
}

size_t bravais::symmetry_size(const std::vector<double>& symm) const {
    if (no_symmetry(symm)) return 1;
    const auto steps = clean_symms(symm);
    size_t r = n_uc_b / steps[0];
    for (size_t i = 1; i < n_dim; i++) r *= n_uc / steps[i];
    return r;
}

// ------------------------------------------------------------ honeycomb
honeycomb::honeycomb(size_t n_uc_, int n_uc_b_, std::vector<double> symmetry)
    : bravais(n_uc_, 2, 2, 3, n_uc_b_ == -1 ? n_uc_ : (size_t)n_uc_b_, 0, std::move(symmetry)) {
    build_bonds();
}

// neighbour order is x-bond, y-bond, z-bond (honeycomb.cpp:36-45)
std::vector<size_t> honeycomb::nns(size_t site) const {
    const size_t c = uc_idx(site);
    if (b_idx(site) == 0) return {idx(c, 1), idx(down(c), 1), idx(down(c, 1), 1)};
    return {idx(c, 0), idx(up(c), 0), idx(up(c, 1), 0)};
}

std::vector<std::vector<size_t>> honeycomb::construct_uc_symmetry(const std::vector<double>& symm) const {
    if (symm.size() == 1 && std::abs(symm[0] - 0.6) < 1e-10)
        throw std::runtime_error("hexagonal 0.6 symmetry is not supported");
    return bravais::construct_uc_symmetry(symm);
}

// symmetry 0.5: every translation, each also composed with the 180-degree
// rotation site -> N-1-site (honeycomb.cpp:84-134)
std::vector<perm_t> honeycomb::construct_symmetry(const std::vector<double>& symm) const {
    if (!half(symm)) return bravais::construct_symmetry(symm);
    std::vector<perm_t> out;
    for (const auto& map : construct_uc_symmetry(symm)) {
        perm_t p(n_total), q(n_total);
        for (size_t i = 0; i < n_total; i++) {
            p[i] = idx(map[uc_idx(i)], b_idx(i));
            q[i] = n_total - 1 - p[i];
        }
        out.push_back(p);
        out.push_back(q);
    }
    return out;
}

std::vector<size_t> honeycomb::construct_symm_basis(const std::vector<double>& symm) const {
    if (half(symm)) return {0};
    return bravais::construct_symm_basis(symm);
}

size_t honeycomb::symmetry_size(const std::vector<double>& symm) const {
    if (half(symm)) return n_total;
    return bravais::symmetry_size(symm);
}

std::vector<std::vector<size_t>> honeycomb::get_hexagons() const {
    std::vector<std::vector<size_t>> out;
    for (size_t c = 0; c < n_total_uc; c++) {
        const size_t r = up(c), rd = down(up(c), 1);
        out.push_back({idx(down(c, 1), 1), idx(c, 0), idx(c, 1), idx(r, 0), idx(rd, 1), idx(rd, 0)});
    }
    return out;
}

// ---------------------------------------------------------------- toric
toric_lattice::toric_lattice(size_t n_uc_) : bravais(n_uc_, 2, 2, 4) { build_bonds(); }

std::vector<size_t> toric_lattice::nns(size_t site) const {
    const size_t c = uc_idx(site);
    if (b_idx(site) == 0) return {idx(c, 1), idx(up(c), 1), idx(down(c, 1), 1), idx(up(down(c, 1)), 1)};
    return {idx(c, 0), idx(down(c), 0), idx(up(c, 1), 0), idx(down(up(c, 1)), 0)};
}

std::vector<plaq> toric_lattice::construct_plaqs() const {
    std::vector<plaq> out;
    for (size_t c = 0; c < n_total_uc; c++) {
        out.push_back({{idx(c, 0), idx(c, 1), idx(up(c, 1), 0), idx(up(c), 1)}, 0});
        out.push_back({{idx(c, 0), idx(c, 1), idx(down(c), 0), idx(down(c, 1), 1)}, 1});
    }
    return out;
}

}  // namespace lattice
