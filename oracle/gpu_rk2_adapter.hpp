// The reference-side drop-in a maintainer adds beside src/fluid/rk2.hpp to
// run integ::rk2integrator::step (src/fluid/solver/integrator.cpp:262-275)
// on the GPU while the reference keeps owning its grid::amrtree ("mode B"
// of the integration notes).  It uses the reference's public data members
// only (grid.hpp:27-120, cell.hpp, prim.hpp, cons.hpp) and the C ABI of
// libbhfluid.so (include/bhfluid.h).  Compiled and exercised here by
// oracle/Makefile (oracle/_ref/libbhadapter.so) and
// tests/test_gpu_parity.py::test_mode_b_adapter_steps_the_reference_tree.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "bhfluid.h"
#include "grid.hpp"

namespace integ {

class gpu_rk2integrator {
public:
    double cfl;
    // the domain and root count the tree was built with (amrtree's
    // constructor arguments, src/fluid/main.cpp:61-89; amrtree keeps none)
    gpu_rk2integrator(double cfl_in, const double dom_l[3], const double dom_h[3], int nqlt) : cfl(cfl_in), nqlt_(nqlt) {
        for (int e = 0; e < 3; e++) { lo_[e] = dom_l[e]; hi_[e] = dom_h[e]; }
    }
    ~gpu_rk2integrator() {
        if (t_) bh_tree_destroy(t_);
    }
    gpu_rk2integrator(const gpu_rk2integrator&) = delete;
    gpu_rk2integrator& operator=(const gpu_rk2integrator&) = delete;

    // Same contract as rk2integrator::step: advances the tree by one step
    // and returns dt; std::domain_error where the reference's equation of
    // state would have thrown (after the state is copied back).
    double step(grid::amrtree& tree) {
        bind(tree);
        pack(tree);
        double dt = 0.0;
        const int rc = bh_rk2_step(t_, cfl, &dt);
        if (rc < 0) throw std::runtime_error(bh_last_error(t_));
        unpack(tree);
        if (rc == BH_EDOMAIN)
            throw std::domain_error("Thermodynamic calculations threw back an error: the density is not positive!");
        return dt;
    }

private:
    static constexpr int NC = grid::cpn * grid::cpn * grid::cpn;               // 8000 stored cells
    static constexpr int NF = (grid::cpn + 1) * grid::cpn * grid::cpn;         // 8400 faces per direction
    bh_tree* t_ = nullptr;
    double lo_[3], hi_[3];
    int nqlt_;
    std::vector<int> meta_;
    std::vector<double> U_, W_, Bf_;

    // block table of the reference's quilt in bh_tree_get_blocks' layout
    static std::vector<int> table(const grid::amrtree& tree) {
        std::unordered_map<const grid::patch*, int> at;
        for (size_t b = 0; b < tree.quilt.size(); b++) at[tree.quilt[b].get()] = (int)b;
        auto idx = [&](const grid::patch* p) { return p ? at.at(p) : -1; };
        std::vector<int> m(tree.quilt.size() * 14);
        for (size_t b = 0; b < tree.quilt.size(); b++) {
            const grid::patch& p = *tree.quilt[b];
            int* x = &m[b * 14];
            x[0] = p.lvl; x[1] = p.ijk[0]; x[2] = p.ijk[1]; x[3] = p.ijk[2];
            x[4] = p.leaf ? 1 : 0;
            x[5] = idx(p.parent);
            for (int h = 0; h < 8; h++) x[6 + h] = idx(p.children[h]);
        }
        return m;
    }
    // the device tree takes the reference's block table (again after a regrid)
    void bind(const grid::amrtree& tree) {
        std::vector<int> m = table(tree);
        if (t_ && m == meta_) return;
        if (!t_ && bh_tree_create(lo_, hi_, nqlt_, tree.mtr.M, tree.mtr.a, tree.mtr.Q, tree.stt.gamma, &t_))
            throw std::runtime_error(bh_last_error(nullptr));
        const int nb = (int)tree.quilt.size();
        if (bh_tree_set_blocks(t_, nb, m.data())) throw std::runtime_error(bh_last_error(t_));
        // the boxes are derived, so check them against the reference's edges
        std::vector<int> meta2(m.size());
        std::vector<double> edges((size_t)nb * 3 * 17);
        bh_tree_get_blocks(t_, meta2.data(), edges.data());
        for (int b = 0; b < nb; b++) {
            const grid::patch& p = *tree.quilt[b];
            const double* e = &edges[(size_t)b * 51];
            if (std::memcmp(e, p.xedge.data(), 17 * sizeof(double)) || std::memcmp(e + 17, p.yedge.data(), 17 * sizeof(double)) ||
                std::memcmp(e + 34, p.zedge.data(), 17 * sizeof(double)))
                throw std::runtime_error("gpu_rk2integrator: block " + std::to_string(b) + " has other edges");
        }
        meta_ = std::move(m);
        U_.assign((size_t)nb * 8 * NC, 0.0);
        W_.assign(U_.size(), 0.0);
        Bf_.assign((size_t)nb * 3 * NF, 0.0);
    }
    void pack(const grid::amrtree& tree) {
        for (size_t b = 0; b < tree.quilt.size(); b++) {
            const grid::patch& p = *tree.quilt[b];
            double* u = &U_[b * 8 * NC];
            double* w = &W_[b * 8 * NC];
            double* f = &Bf_[b * 3 * NF];
            for (int n = 0; n < NC; n++) {
                const cons& c = p.cells[n].U;
                const prim& q = p.cells[n].W;
                u[n] = c.D;
                for (int d = 0; d < 3; d++) u[(1 + d) * NC + n] = c.S[d];
                u[4 * NC + n] = c.tau;
                for (int d = 0; d < 3; d++) u[(5 + d) * NC + n] = c.B[d];
                w[n] = q.rho;
                w[NC + n] = q.eps;
                for (int d = 0; d < 3; d++) {
                    w[(2 + d) * NC + n] = q.v[d];
                    w[(5 + d) * NC + n] = q.B[d];
                }
            }
            std::memcpy(f, p.Bfx.data(), NF * sizeof(double));
            std::memcpy(f + NF, p.Bfy.data(), NF * sizeof(double));
            std::memcpy(f + 2 * NF, p.Bfz.data(), NF * sizeof(double));
        }
        if (bh_tree_set_state(t_, U_.data(), W_.data(), Bf_.data())) throw std::runtime_error(bh_last_error(t_));
    }
    // The derived members of prim (p, T, h, lor, b2) are not part of the
    // exchanged state; the reference's step recomputes them where it reads
    // them, and a caller that reads them between steps recomputes them.
    void unpack(grid::amrtree& tree) {
        if (bh_tree_get_state(t_, U_.data(), W_.data(), Bf_.data())) throw std::runtime_error(bh_last_error(t_));
        for (size_t b = 0; b < tree.quilt.size(); b++) {
            grid::patch& p = *tree.quilt[b];
            const double* u = &U_[b * 8 * NC];
            const double* w = &W_[b * 8 * NC];
            const double* f = &Bf_[b * 3 * NF];
            for (int n = 0; n < NC; n++) {
                cons& c = p.cells[n].U;
                prim& q = p.cells[n].W;
                c.D = u[n];
                for (int d = 0; d < 3; d++) c.S[d] = u[(1 + d) * NC + n];
                c.tau = u[4 * NC + n];
                for (int d = 0; d < 3; d++) c.B[d] = u[(5 + d) * NC + n];
                q.rho = w[n];
                q.eps = w[NC + n];
                for (int d = 0; d < 3; d++) {
                    q.v[d] = w[(2 + d) * NC + n];
                    q.B[d] = w[(5 + d) * NC + n];
                }
            }
            std::memcpy(p.Bfx.data(), f, NF * sizeof(double));
            std::memcpy(p.Bfy.data(), f + NF, NF * sizeof(double));
            std::memcpy(p.Bfz.data(), f + 2 * NF, NF * sizeof(double));
        }
    }
};

}  // namespace integ
