// The device-resident system (fc_system of include/fcap.h), shared by the C ABI
// layer and the learned preconditioner.
#pragma once

#include "fcap_gnn.h"
#include "fcap_internal.h"

struct fc_gnn;

struct fc_system {
    fc_context *ctx = nullptr;
    fcap::MeshDev mesh;
    fcap::Structure st;
    fcap::DevBuf ij, ji, kface, kord;
    bool needs_order = false;
    fcap::WallDev walls;
    fcap::OcDev oc;
    int ntotal = 0;
    fcap::DevBuf a, su, fi, mu, phi, su0, vis, flux;
    // staging of host-resident fields and flmass of the pressure-correction assembly
    fcap::DevBuf cp_stage[12], cp_flmass;
    // Gauss gradient: per-cell boundary lists of the last boundary groups, staging of u,
    // the component arrays and the (3, numCells) output
    int grad_bnd[10] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    fcap::DevBuf grad_cb_ptr, grad_cb_q, grad_u, grad_xyz, grad_out;
    // pp over numTotal for bpres + grad(pp), and the outlet factor of the velocity BCs
    fcap::DevBuf pp_total, vbc_fac;
    fcap::GradScratch grad_scratch;
    bool su0_zero = true;
    double last_ms[4] = {0, 0, 0, 0};
    // bumped by every assembly; the network's graph data is rebuilt when stale
    long assembly_version = 0;
    fcap::GnnGraph graph;
    long graph_version = -1;
    fc_gnn *net = nullptr;     // preconditioner of FC_METHOD_FCG_LEARNED
};

namespace fcap {
// Graph data of the system's current matrix (rebuilt after re-assembly).
GnnGraph &system_graph(fc_system *sys);
// The attached network as the preconditioner hook of flexible CG (apply == nullptr if none).
PrecondHook gnn_hook(fc_system *sys);
// the same preconditioner without the network (FC_METHOD_FCG_JACOBI)
PrecondHook jacobi_hook(fc_system *sys);
}  // namespace fcap
