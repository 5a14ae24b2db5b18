// model.hpp -- Hamiltonians of the reference's models as lists of local
// operators (include/model/kitaev.hpp, toric.hpp, abstract_model.hpp).
//
// A local operator acts on k sites as a dense 2^k x 2^k complex matrix
// whose local basis index has bit i = state of acts_on[i] (bit 1 <=>
// sigma = -1), i.e. acts_on[0] is the least significant factor of the
// Kronecker product (abstract_model.cpp:28-33).
#pragma once
#include <array>
#include <complex>
#include <memory>
#include <vector>

#include "lattice.hpp"

namespace model {

using cplx = std::complex<double>;

struct local_op {
    std::vector<size_t> acts_on;
    size_t dim;               // 2^k
    std::vector<cplx> mat;    // row-major dim x dim
};

// Pauli matrices and their Kronecker products
struct dense {
    size_t n;
    std::vector<cplx> a;
};
dense sx();
dense sy();
dense sz();
dense kron(const std::vector<dense>& factors);  // factors[0] least significant

class abstract_model {
   public:
    virtual ~abstract_model() = default;
    lattice::bravais& get_lattice() { return *lattice_; }
    const std::vector<local_op>& get_hamiltonian() const { return hamiltonian_; }
    virtual bool supports_helper_hamiltonian() const { return false; }
    virtual void add_helper_hamiltonian(double) {}
    void remove_helper_hamiltonian() {  // abstract_model.cpp:60-65
        for (size_t i = 0; i < helpers_; i++) hamiltonian_.pop_back();
        helpers_ = 0;
    }

   protected:
    std::unique_ptr<lattice::bravais> lattice_;
    std::vector<local_op> hamiltonian_;
    size_t helpers_ = 0;
};

// Kitaev honeycomb: J_x sx sx, J_y sy sy, J_z sz sz on x/y/z bonds plus an
// optional field -h sx + h sy + h sz per site (kitaev.cpp:29-50)
class kitaev : public abstract_model {
   public:
    kitaev(size_t size, const std::vector<double>& J, int size_b = -1,
           const std::vector<double>& symmetry = {1}, double h = 0);
    bool supports_helper_hamiltonian() const override { return true; }
    // -strength (W_p - 1) per hexagon (kitaev.cpp:52-64)
    void add_helper_hamiltonian(double strength) override;
    // closed-form flux-free ground energy per site (kitaev.cpp:66-100)
    double exact_energy() const;
    lattice::honeycomb& honeycomb_lattice() { return static_cast<lattice::honeycomb&>(*lattice_); }

   private:
    std::array<double, 3> J_;
    double h_;
};

// toric code: J prod sz on plaquettes, J prod sx on vertices (toric.cpp:32-47)
class toric : public abstract_model {
   public:
    toric(size_t size, double J = -1.);
};

// hexagon plaquette operator sx sy sz sx sy sz used by --evaluate (main.cpp:645-647)
dense plaquette_op();

}  // namespace model
