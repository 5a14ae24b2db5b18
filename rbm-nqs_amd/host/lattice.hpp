// lattice.hpp -- periodic Bravais lattices of the reference's models
// (include/lattice/bravais.hpp, honeycomb.hpp, toric_lattice.hpp): site
// indexing, neighbour bonds, hexagons / plaquettes and the symmetry
// permutations used by the symmetric RBM and the Pfaffian index map.
#pragma once
#include <array>
#include <cstddef>
#include <string>
#include <vector>

namespace lattice {

struct bond {
    size_t a, b, type;
};

// permutation table: perm[i] is Eigen's indices()(i) of the reference
using perm_t = std::vector<size_t>;

class bravais {
   public:
    const size_t n_uc, n_dim, n_basis, n_coordination, n_total_uc, n_total, n_uc_b, h_shift;

    bravais(size_t n_uc, size_t n_dim, size_t n_basis, size_t n_coordination, size_t n_uc_b = 0,
            size_t h_shift = 0, std::vector<double> default_symmetry = {1});
    virtual ~bravais() = default;

    size_t uc_idx(size_t site) const { return site / n_basis; }
    size_t b_idx(size_t site) const { return site % n_basis; }
    size_t idx(size_t uc, size_t b) const { return uc * n_basis + b; }
    // move a unit cell by `step` cells along `dir` (bravais.cpp:66-105)
    size_t up(size_t uc, size_t dir = 0, size_t step = 1) const;
    size_t down(size_t uc, size_t dir = 0, size_t step = 1) const;

    virtual std::vector<size_t> nns(size_t site) const = 0;
    const std::vector<bond>& get_bonds() const { return bonds_; }

    virtual std::vector<perm_t> construct_symmetry(const std::vector<double>& symm) const;
    std::vector<perm_t> construct_symmetry() const { return construct_symmetry(default_symmetry_); }
    virtual std::vector<size_t> construct_symm_basis(const std::vector<double>& symm) const;
    std::vector<size_t> construct_symm_basis() const { return construct_symm_basis(default_symmetry_); }
    virtual std::vector<std::vector<size_t>> construct_uc_symmetry(const std::vector<double>& symm) const;
    virtual size_t symmetry_size(const std::vector<double>& symm) const;
    size_t symmetry_size() const { return symmetry_size(default_symmetry_); }
    const std::vector<double>& default_symmetry() const { return default_symmetry_; }

   protected:
    std::vector<bond> bonds_;
    std::vector<double> default_symmetry_;
    void build_bonds();  // neighbours of every basis-0 site, type = slot
    std::vector<size_t> clean_symms(const std::vector<double>& symm) const;
    static bool no_symmetry(const std::vector<double>& s) { return s.empty() || s[0] == 0; }
};

class honeycomb : public bravais {
   public:
    explicit honeycomb(size_t n_uc, int n_uc_b = -1, std::vector<double> symmetry = {1});
    std::vector<size_t> nns(size_t site) const override;
    std::vector<perm_t> construct_symmetry(const std::vector<double>& symm) const override;
    std::vector<size_t> construct_symm_basis(const std::vector<double>& symm) const override;
    std::vector<std::vector<size_t>> construct_uc_symmetry(const std::vector<double>& symm) const override;
    size_t symmetry_size(const std::vector<double>& symm) const override;
    // 6-site plaquettes, honeycomb.cpp:201-209
    std::vector<std::vector<size_t>> get_hexagons() const;
    using bravais::construct_symm_basis;
    using bravais::construct_symmetry;
    using bravais::symmetry_size;

   private:
    static bool half(const std::vector<double>& s) { return s.size() == 1 && s[0] == 0.5; }
};

struct plaq {
    std::array<size_t, 4> idxs;
    size_t type;  // 0 plaquette (sigma^z), 1 vertex (sigma^x)
};

class toric_lattice : public bravais {
   public:
    explicit toric_lattice(size_t n_uc);
    std::vector<size_t> nns(size_t site) const override;
    std::vector<plaq> construct_plaqs() const;  // toric_lattice.cpp:57-69
};

}  // namespace lattice
