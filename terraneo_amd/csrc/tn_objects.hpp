// Opaque objects behind the C ABI handles.
#pragma once

#include "tn_internal.hpp"

struct tn_eps {
    const tn_mesh* mesh = nullptr;
    int dtype = TN_F64;
    const void* eta = nullptr;       // nodal viscosity [10][n][n][nr]
    const void* eta_cell = nullptr;  // per hex cell viscosity [10][n-1][n-1][nr-1]; overrides eta when set
    int bc_cmb = TN_NEUMANN, bc_surf = TN_NEUMANN;
    int apply_mode = TN_REPLACE, comm_mode = TN_COMM_ADDITIVE;
    bool diagonal = false;
    double penalty_eps = 1.0;
    // 0: the penalty is added at every node copy before the interface sum (the
    // reference's rule); 1: at the owned copy only (once per node, symmetric)
    int penalty_once = 0;
    // fs/fs rotation penalty: y += eps * (G^{-1} beta) x p, beta = sum_owned p x u
    tn::DevBuf<double> pen;  // [0..3) beta, [4..13) S = eps G^{-1}
    bool penalty_ready = false;
    // boundary-node quadratic forms of the matrix-free march kernel, rebuilt
    // after tn_eps_set_eta / tn_eps_set_boundary_conditions
    tn::DevBuf<unsigned char> blk;
    bool blk_ready = false;
    // stored local matrices (Galerkin coarse operators), packed symmetric
    int stored_dtype = -1;          // -1: matrix-free
    tn::DevBuf<unsigned char> stored;  // tn::stored_at layout, stored_dtype elements
    // stored apply: per-cell corner sums between its two passes (12 per cell
    // plus 12 per radial chunk top), allocated on first use
    tn::DevBuf<unsigned char> cellsum;
    // one 18x18 local matrix in transit (get / set_local_matrix), allocated on first use
    tn::DevBuf<double> lm;
    int64_t cells() const
    {
        const auto& m = mesh->lvl;
        return 10LL * (m.n - 1) * (m.n - 1) * (m.nr - 1);
    }
};

struct tn_stokes {
    const tn_mesh* vel = nullptr;
    const tn_mesh* pre = nullptr;
    tn_eps* A = nullptr;  // viscous block on the velocity mesh; its BCs govern B and B^T
    tn::DevBuf<unsigned char> rhs_tmp;  // nodal buoyancy before the mass matrix (tn_stokes_rhs_from_temperature)
    tn::DevBuf<unsigned char> div_tmp;  // B's per-cell corner sums (div_scratch_elems)
};

namespace tn {
void stokes_apply(tn_stokes* K, const void* u, const void* p, void* yu, void* yp, hipStream_t s);
template <typename T>
void grad_t(tn_stokes* K, const void* p, void* u, bool add, bool comm, hipStream_t s);
// C++ entry used by the Stokes and solver layers
void eps_apply(tn_eps* op, const void* src, void* dst, hipStream_t s);
}
