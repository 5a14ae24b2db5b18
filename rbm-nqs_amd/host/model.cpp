// model.cpp -- see model.hpp.
#include "model.hpp"

#include <cmath>

namespace model {

dense sx() { return {2, {0., 1., 1., 0.}}; }
dense sy() { return {2, {0., cplx(0, -1), cplx(0, 1), 0.}}; }
dense sz() { return {2, {1., 0., 0., -1.}}; }

dense kron(const std::vector<dense>& factors) {
    dense out{1, {1.}};
    for (const auto& f : factors) {  // out <- f (x) out
        dense r{f.n * out.n, std::vector<cplx>(f.n * out.n * f.n * out.n)};
        for (size_t i = 0; i < f.n; i++)
            for (size_t j = 0; j < f.n; j++) {
                const cplx v = f.a[i * f.n + j];
                if (v == 0.) continue;
                for (size_t k = 0; k < out.n; k++)
                    for (size_t l = 0; l < out.n; l++)
                        r.a[(i * out.n + k) * r.n + j * out.n + l] = v * out.a[k * out.n + l];
            }
        out = std::move(r);
    }
    return out;
}

static dense scaled(dense d, cplx s) {
    for (auto& x : d.a) x *= s;
    return d;
}

dense plaquette_op() { return kron({sx(), sy(), sz(), sx(), sy(), sz()}); }

kitaev::kitaev(size_t size, const std::vector<double>& J, int size_b, const std::vector<double>& symmetry,
               double h)
    : h_(h) {
    for (int i = 0; i < 3; i++) J_[i] = J[i % J.size()];
    lattice_ = std::make_unique<lattice::honeycomb>(size, size_b, symmetry);
    const dense bond_ops[3] = {scaled(kron({sx(), sx()}), J_[0]), scaled(kron({sy(), sy()}), J_[1]),
                               scaled(kron({sz(), sz()}), J_[2])};
    for (const auto& b : lattice_->get_bonds())
        hamiltonian_.push_back({{b.a, b.b}, 4, bond_ops[b.type].a});
    if (std::abs(h) > 1e-10) {
        dense x = sx(), y = sy(), z = sz();
        std::vector<cplx> op(4);
        for (int i = 0; i < 4; i++) op[i] = -h * x.a[i] + h * y.a[i] + h * z.a[i];
        for (size_t i = 0; i < lattice_->n_total; i++) hamiltonian_.push_back({{i}, 2, op});
    }
}

void kitaev::add_helper_hamiltonian(double strength) {
    dense p = plaquette_op();
    for (size_t i = 0; i < p.n; i++) p.a[i * p.n + i] -= 1.;
    p = scaled(p, -strength);
    const auto hex = honeycomb_lattice().get_hexagons();
    helpers_ = hex.size();
    for (const auto& hx : hex) hamiltonian_.push_back({hx, p.n, p.a});
}

double kitaev::exact_energy() const {
    // |f(q)| summed over the Brillouin zone for the four periodic /
    // antiperiodic boundary conditions; the largest sum gives the ground
    // energy -sum/(2N)
    const double na = (double)lattice_->n_uc_b, nb = (double)lattice_->n_uc;
    double best = -1e10;
    for (int c1i = 0; c1i < 2; c1i++)
        for (int c2i = 0; c2i < 2; c2i++) {
            const double c1 = 0.5 / nb * c1i, c2 = 0.5 / na * c2i;
            double sum = 0;
            for (size_t i = 0; i < lattice_->n_uc; i++)
                for (size_t j = 0; j < lattice_->n_uc_b; j++) {
                    const double p1 = i / nb + c1, p2 = j / na + c2;
                    const cplx f = 2. * (J_[0] * std::exp(cplx(0, 2 * M_PI * p1)) +
                                         J_[1] * std::exp(cplx(0, 2 * M_PI * p2)) + J_[2]);
                    sum += std::abs(f);
                }
            best = std::max(best, sum);
        }
    return -0.5 * best / (double)lattice_->n_total;
}

toric::toric(size_t size, double J) {
    auto* lat = new lattice::toric_lattice(size);
    lattice_.reset(lat);
    const dense plaq = scaled(kron({sz(), sz(), sz(), sz()}), J);
    const dense vert = scaled(kron({sx(), sx(), sx(), sx()}), J);
    for (const auto& p : lat->construct_plaqs()) {
        std::vector<size_t> s(p.idxs.begin(), p.idxs.end());
        hamiltonian_.push_back({s, 16, p.type == 0 ? plaq.a : vert.a});
    }
}

}  // namespace model
