// mist_internal.h -- host-side context of libmist_hip.so and the launchers each .hip file
// provides. Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mist_hip.h"
#include "mist_device.h"

#define MIST_HIP_TRY(expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
            return MIST_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

namespace mist {
struct MGState;
void mg_free(MGState* s);
struct FasState;
void fas_free(FasState* s);
}

enum { MIST_STAGE_NORMAL = 0, MIST_STAGE_TRANSPORT, MIST_STAGE_PLIC, MIST_STAGE_BAND, MIST_STAGE_MG, MIST_STAGE_COUNT };

// Per-kernel timing with HIP events recorded on the context's stream around each launch of
// the profiled kernels (enabled by mist_profile_enable); elapsed times are collected when
// the profile is read. Work units per launch (interface cells for ELVIRA, band cells for
// the cut, cells for the rest) are accumulated on the device.
struct KernelProfile {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[MIST_PROF_COUNT];
    double ms[MIST_PROF_COUNT] = {};
    int64_t launches[MIST_PROF_COUNT] = {};
    int64_t host_units[MIST_PROF_COUNT] = {};
    unsigned long long* d_units = nullptr;  // device accumulators, one per kernel
};

struct mist_ctx {
    mist_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    mist::DevMesh m{};
    mist::DevParams dp{};
    std::vector<double> hx, hy, hz, hxm, hym, hzm;  // host copies, index q = reference i - imino
    double* geom = nullptr;
    double* trig = nullptr;      // sin(pi c), sin(2 pi c), cos(pi c) of the coordinates
    mist::VelTab vt{};
    void* f[MIST_FIELD_COUNT] = {};
    int64_t flo[MIST_FIELD_COUNT][3] = {};
    int64_t fn[MIST_FIELD_COUNT][3] = {};
    int64_t felem[MIST_FIELD_COUNT] = {};
    // work buffers
    double* pv = nullptr;        // projected cell vertices, (Nx+1)(Ny+1)(Nz+1) x 3
    int* band_list = nullptr;    // interior cells with |band| <= nband (linear cell index)
    int* iface_list = nullptr;   // interior interface cells (ELVIRA); cells the main cut pass defers
    int* defer_list = nullptr;   // cells for the general cut walk
    void* cut_list = nullptr;    // CutCell entries of the band cells to cut
    int* vtx_list = nullptr;     // vertices of band cells to back-trace (linear vertex index)
    bool vtx_list_ok = false;    // vtx_list matches the band (set only by calls that reuse the list)
    int* elv_ties = nullptr;     // ELVIRA's near-tied cells: (interface-list index, candidate mask)
    int* counters = nullptr;     // [0] band cells, [1] interface cells, [2] overflow flag, [4] cells
                                 // deferred by the main cut pass, [5] its work queue, [6] cells for
                                 // the general walk, [7] the wide pass's work queue, [8] cells to
                                 // cut, [9] vertices to back-trace, [10] ELVIRA's near-tied cells,
                                 // [11] ELVIRA's work queue
    double* red = nullptr;       // reduction partials
    int red_blocks = 0;
    int* h_counters = nullptr;   // pinned
    int tree_blocks = 0;         // resident blocks of the main cut pass on the device
    int tree_cus = 0;            // compute units of the device
    int tree_blocks_mom = 0;     // resident blocks of the main cut pass with momentum sums
    int tree_blocks_adj = 0;     // resident blocks of the reverse cut pass
    int elv_blocks = 0;          // resident blocks of ELVIRA's scoring kernel
    double* ns_sums = nullptr;   // 6 sums per interior cell of the momentum predictor's cut
    double* h_red = nullptr;     // pinned
    // mist_vof_step with deferred checks: the velocity maxima the next step's dt needs are
    // reduced right after the velocity is defined, copied to pinned memory and marked by an
    // event, so the next step waits for that event instead of for the whole stream. They are
    // used only when no other call came in between (api_seq counts the entry points).
    uint64_t api_seq = 0;
    uint64_t vmax_seq = 0;
    bool vmax_pending = false;
    double* h_vmax = nullptr;    // pinned, 3 doubles
    hipEvent_t vmax_ev = nullptr;
    uint64_t vf_version = 1, normals_version = 0, band_version = 0;
    bool defer_checks = false;   // transport / VJP calls skip their host round trip (mist_set_deferred_checks)
    bool band_list_valid = false;
    mist::MGState* mg = nullptr;
    mist::FasState* fas = nullptr;  // SL pressure: multigrid levels (child contexts) and Anderson state
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double stage_ms[MIST_STAGE_COUNT] = {};
    bool timing = false;
    KernelProfile prof;
    double bc_wall[6][3] = {};   // wall velocity of sides x-, x+, y-, y+, z-, z+ (reflect BC)
    int* elv_best = nullptr;     // ELVIRA candidate kept per interior cell (-1: none)
    double* elv_dist = nullptr;  // computeDist of the 27 candidates per interface cell (allocated on first use)
    double* adj_lam = nullptr;   // adjoint work: dL/dVF of the step input (VF extents)
    double* adj_plane = nullptr; // adjoint work: planes nx, ny, nz, D (4 cell-extent arrays)
    double* adj_pv = nullptr;    // adjoint work: back-traced corners (3 per grid vertex)
    double* adj_sums = nullptr;  // adjoint work: 6 cut sums per interior cell (vol, vLiq used)
    uint64_t plic_version = 0;   // VF version the PLIC constants were computed for
    hipGraph_t vjp_graph = nullptr;     // the reverse step's launches, captured once per dt
    hipGraphExec_t vjp_exec = nullptr;
    double vjp_dt = 0.0;
    std::vector<double*> ckpt;   // VF checkpoints (device)
    std::string err;
};

namespace mist {
// profiling helpers (mist_capi.hip)
hipEvent_t prof_begin(mist_ctx* ctx, int kernel);
void prof_end(mist_ctx* ctx, int kernel, hipEvent_t a, const int* device_units, int64_t host_units);
// mist_vof.hip
int launch_trig_tables(mist_ctx* ctx, int64_t n);
int launch_define_velocity(mist_ctx* ctx, double t);
int launch_define_velocity_vmax(mist_ctx* ctx, double t);
int launch_maxabs_uvw(mist_ctx* ctx, double out[3]);
int launch_neumann(mist_ctx* ctx, int field);
int launch_neumann_cells(mist_ctx* ctx, double* A);  // Neumann! of a cell-extent array of ctx's mesh
int launch_clear_vof_outer(mist_ctx* ctx);           // VF outside the cell extents = 0
int launch_borders(mist_ctx* ctx, int field);
int launch_band(mist_ctx* ctx);
int launch_normals(mist_ctx* ctx);
int launch_plic(mist_ctx* ctx);
// reuse_list: keep the band cells' vertex list of the previous call when vtx_list_ok says the
// band has not changed since (the SL pressure's repeated A! evaluations on one band)
int launch_project_vertices(mist_ctx* ctx, double dt, bool reuse_list = false, bool count_cleared = false);
int launch_transport(mist_ctx* ctx, double dt);
int launch_tree_momentum(mist_ctx* ctx, bool corr, double dt);
int launch_cut_list(mist_ctx* ctx);
int launch_init_vof(mist_ctx* ctx, int shape, const double* args);
int launch_props(mist_ctx* ctx);
int launch_pressure_rhs(mist_ctx* ctx);
int launch_corrector(mist_ctx* ctx, double dt);
int launch_reduce(mist_ctx* ctx, int kind, int field, double* out);
int launch_plic_dist_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* vf, double* out, int64_t m);
int launch_plic_vf_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* dist, double* out, int64_t m);
// mist_ns.hip
int launch_ns_advect(mist_ctx* ctx, double dt);
int launch_ns_forces(mist_ctx* ctx, double dt);
int launch_curvature(mist_ctx* ctx);
int launch_surface_force(mist_ctx* ctx);
int launch_reflect_bc(mist_ctx* ctx, int fu, int fv, int fw);
int launch_interp_face(mist_ctx* ctx, int fu, int fv, int fw);
int launch_interp_center(mist_ctx* ctx);
int launch_divergence(mist_ctx* ctx, double dt, bool reuse_vertex_list = false);
// mist_adjoint.hip
int launch_transport_vjp(mist_ctx* ctx, double dt);
int vjp_prepare(mist_ctx* ctx);  // what the reverse step's first call sets up on the host
// mist_mg.hip
int pressure_solve(mist_ctx* ctx, double dt, double tol, int precision, int64_t* cycles, double* resid);
// mist_fas.hip
int pressure_solve_sl(mist_ctx* ctx, double dt, double tol, int64_t* cycles, double* resid, int64_t* smooth_iters);
int pressure_operator_sl(mist_ctx* ctx, double dt);
int pressure_jacobian_sl(mist_ctx* ctx, double dt, double* out);
}  // namespace mist
