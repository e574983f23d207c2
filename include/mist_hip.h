/*
 * mist_hip.h -- C ABI of libmist_hip.so, the MI355X (gfx950) implementation of the
 * per-step hot path of Mist.jl / NS_Multiphase:
 *   - semi-Lagrangian geometric VOF transport (ELVIRA normals, PLIC plane constants,
 *     back-traced cell pre-images split into 5 tets, tet/mesh-plane/PLIC cutting), and
 *   - the variable-density pressure Poisson multigrid solve (red-black Gauss-Seidel,
 *     residual, 8-cell restriction, trilinear prolongation) in FP64 or as FP32 V-cycles
 *     inside an FP64 defect-correction loop.
 *
 * Conventions
 *   - Plain C: no C++ types, no torch types. Every entry point returns an int status
 *     (MIST_OK = 0); mist_last_error() gives the message of the last failure.
 *   - Fields live in device memory inside a context. Host arrays cross the boundary by
 *     mist_upload / mist_download as flat copies of the reference's ghosted, column-major
 *     (i fastest) OffsetArray parents, so a Julia caller passes `parent(A)` unchanged.
 *   - Indices follow the reference: interior cells are 1..N per axis, ghosts extend to
 *     1-nghost and N+nghost, nghost = 3 + max(0, ceil(CFL)-2) (src/Mesh.jl:66). The volume
 *     fraction carries 3 extra ghost layers (src/Tools.jl:144); face density/viscosity
 *     arrays carry one extra face on their own axis (src/Tools.jl:169-177).
 *   - All work of one context is ordered on one HIP stream. One host thread per context.
 *     download, reduce and the calls that return scalars synchronise that stream.
 */
#ifndef MIST_HIP_H
#define MIST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------------- */
enum {
    MIST_OK = 0,
    MIST_ERR_PARAM = 1,     /* invalid parameter / unsupported option */
    MIST_ERR_OOM = 2,       /* device allocation failed */
    MIST_ERR_HIP = 3,       /* HIP runtime error */
    MIST_ERR_NOCONV = 4,    /* solver stopped at its cycle limit above tolerance */
    MIST_ERR_STATE = 5,     /* call order / missing field */
    MIST_ERR_OVERFLOW = 6,  /* a geometric work bound was exceeded (cut depth, walk) */
    MIST_ERR_NODEVICE = 7   /* no HIP device available */
};

/* ---- string-valued options of the reference `parameters` struct, as enums -------- */
enum { MIST_PS_SEMI_LAGRANGIAN = 0, MIST_PS_FINITE_DIFFERENCE = 1 };      /* pressure_scheme */
enum { MIST_NORMAL_ELVIRA = 0 };                                           /* normalMethod */
enum { MIST_VEL_NONE = 0, MIST_VEL_DEFORMATION = 1, MIST_VEL_DEFORMATION3D = 2,
       MIST_VEL_ROTATION = 3, MIST_VEL_DIVFLOW = 4, MIST_VEL_DIVFLOW3D = 5,
       MIST_VEL_AIRBLAST = 6 };                                             /* VFVelocity */
enum { MIST_PROJ_RK4 = 0, MIST_PROJ_EULER = 1, MIST_PROJ_MIDPOINT = 2 };    /* projection_method */
enum { MIST_TESS_5_TETS = 0 };                                              /* tesselation */
enum { MIST_CASE_GENERIC = 0, MIST_CASE_ZALESAK = 1 };                      /* test_case */
enum { MIST_MG_FP64 = 0, MIST_MG_MIXED = 1 };                               /* new: V-cycle precision */
/* new: coarse operator. MIST_COARSE_FACE_AVERAGE: coarse face coefficients are the mean of
 * the fine ones. MIST_COARSE_RESTRICTED_VF: the reference's coarse densities, compute_dens!
 * of the 8-cell mean of the finer level's VF (src/Multigrid.jl:725-737); its V-cycle can
 * diverge at high density ratios (1000:1), where the solve then reports MIST_ERR_NOCONV. */
enum { MIST_COARSE_FACE_AVERAGE = 0, MIST_COARSE_RESTRICTED_VF = 1 };

/*
 * POD mirror of the fields of `parameters` (src/Parameters.jl:3-91) that the hot path
 * reads. Strings become the enums above. Fields marked "new" have no reference
 * counterpart; mist_params_default() fills them.
 */
typedef struct mist_params {
    double mu_liq, mu_gas, rho_liq, rho_gas, sigma;
    double grav_x, grav_y, grav_z;
    double Lx, Ly, Lz;
    double tFinal;
    int64_t Nx, Ny, Nz;
    int64_t stepMax;
    double max_dt;          /* reference default Inf */
    double CFL;
    int64_t xper, yper, zper;
    int64_t solveNS;
    double tol;
    int64_t pressure_scheme;
    int64_t mg_lvl;
    double VFlo, VFhi;      /* defaults 1e-10 and 1-VFlo */
    int64_t nband;          /* default ceil(CFL) */
    int64_t normal_method;
    int64_t vf_velocity;
    int64_t projection_method;
    int64_t tesselation;
    int64_t test_case;
    /* new */
    double vel_period;      /* T in cos(pi t / T) of the deformation fields; 2.0 = reference */
    int64_t mg_precision;   /* MIST_MG_FP64 or MIST_MG_MIXED */
    int64_t mg_nu1, mg_nu2; /* smoothing sweeps, reference 10 and 10 */
    int64_t mg_max_cycles;  /* outer cycle limit, reference 10000 */
    int64_t mg_coarse_op;   /* MIST_COARSE_FACE_AVERAGE (default) or MIST_COARSE_RESTRICTED_VF */
    int64_t mg_coarse_sweeps; /* RB-GS sweeps on the coarsest level */
} mist_params;

/* ---- field identifiers (device-resident arrays of a context) --------------------- */
enum {
    MIST_FIELD_VOF = 0,   /* VF: volume fraction, cell extents + 3 extra ghost layers, f64 */
    MIST_FIELD_NX = 1, MIST_FIELD_NY = 2, MIST_FIELD_NZ = 3,   /* interface normal, f64 */
    MIST_FIELD_PLIC_D = 4,  /* D: PLIC plane constant, n.x = D, f64 */
    MIST_FIELD_BAND = 5,    /* band: int16 */
    MIST_FIELD_U = 6, MIST_FIELD_V = 7, MIST_FIELD_W = 8,      /* cell-centre velocity */
    MIST_FIELD_UF = 9, MIST_FIELD_VF = 10, MIST_FIELD_WF = 11, /* face velocity, cell extents */
    MIST_FIELD_P = 12,      /* pressure */
    MIST_FIELD_DIVG = 13,   /* divergence diagnostic */
    MIST_FIELD_DENX = 14, MIST_FIELD_DENY = 15, MIST_FIELD_DENZ = 16, /* face density */
    MIST_FIELD_RHS = 17,    /* pressure right-hand side b (interior cells are used) */
    MIST_FIELD_US = 18, MIST_FIELD_VS = 19, MIST_FIELD_WS = 20, /* predicted cell velocity */
    MIST_FIELD_VISCX = 21, MIST_FIELD_VISCY = 22, MIST_FIELD_VISCZ = 23, /* face viscosity */
    MIST_FIELD_CURV = 24,   /* Curve: interface curvature (height function) */
    MIST_FIELD_SFX = 25, MIST_FIELD_SFY = 26, MIST_FIELD_SFZ = 27, /* face surface force */
    MIST_FIELD_MASK = 28,   /* mask: int16, bit d set = mask[i,j,k][d+1] of the reference */
    /* adjoints (mist_transport_vjp): of VF (VF extents) and of the face velocities */
    MIST_FIELD_VOF_BAR = 29, MIST_FIELD_UF_BAR = 30, MIST_FIELD_VFACE_BAR = 31, MIST_FIELD_WF_BAR = 32,
    MIST_FIELD_COUNT = 33
};

enum { MIST_REDUCE_SUM = 0, MIST_REDUCE_MAX = 1, MIST_REDUCE_MIN = 2, MIST_REDUCE_MAXABS = 3 };

enum { MIST_IC_CIRCLE = 0, MIST_IC_SPHERE = 1, MIST_IC_BUBBLE3D = 2 };

typedef struct mist_ctx mist_ctx;

/* Fill *p with the reference defaults (src/Parameters.jl:3-91) plus the new fields. */
int mist_params_default(mist_params* p);

/* Create a context on HIP device `device`: builds the mesh (src/Mesh.jl:55-196, one rank)
 * and allocates every field in device memory (src/Tools.jl:136-191). */
int mist_ctx_create(const mist_params* p, int device, mist_ctx** out);
int mist_ctx_destroy(mist_ctx* ctx);
const char* mist_last_error(const mist_ctx* ctx);   /* ctx may be NULL: last global error */

/* Extents of a field: lo = lowest (1-based reference) index per axis, n = length per axis,
 * elem = element size in bytes. */
int mist_field_extents(const mist_ctx* ctx, int field, int64_t lo[3], int64_t n[3], int64_t* elem);
/* Device pointer of a field (for zero-copy interop; owned by the context). Fetching it marks
 * what the field feeds as stale (as mist_field_modified does); a caller that writes through
 * the pointer after later stage calls must call mist_field_modified after the write. */
int mist_field_device_ptr(mist_ctx* ctx, int field, void** dptr);
/* Declare that `field` was written through its device pointer: the normals, PLIC constants
 * and band list derived from it are recomputed when next needed (VF: all three; a normal
 * component: normals and constants; D: constants; band: the band-cell list). No reference
 * counterpart (the reference's arrays carry no derived state). */
int mist_field_modified(mist_ctx* ctx, int field);

int mist_upload(mist_ctx* ctx, int field, const void* host, size_t nbytes);
int mist_download(mist_ctx* ctx, int field, void* host, size_t nbytes);

/* Initial volume fraction over the cell extents (IC! loops over imino_:imaxo_):
 * MIST_IC_CIRCLE args = {rad, xo, yo}          (VFcircle, src/Tools.jl:790-902)
 * MIST_IC_SPHERE args = {rad, xo, yo, zo}      (VFsphere, src/Tools.jl:907-922)
 * MIST_IC_BUBBLE3D args = {rad, xo, yo, zo}    (VFbubble3d, src/Tools.jl:1061-1076) */
int mist_init_vof(mist_ctx* ctx, int shape, const double* args, int nargs);

/* Stage entry points (one each per reference function). */
int mist_neumann(mist_ctx* ctx, int field);             /* Neumann!  src/Tools.jl:4-18 */
int mist_update_borders(mist_ctx* ctx, int field);      /* update_borders!  src/Parallel.jl:63-143 */
int mist_define_velocity(mist_ctx* ctx, double t);      /* defineVelocity!  src/Tools.jl:648-693 */
int mist_compute_dt(mist_ctx* ctx, double* dt);         /* compute_dt  src/Tools.jl:317-337 */
int mist_compute_band(mist_ctx* ctx);                   /* computeBand!  src/VF.jl:16-74 */
int mist_compute_normal(mist_ctx* ctx);                 /* computeNormal!/ELVIRA!  src/VF.jl:79, src/ELVIRA.jl:6-137 */
int mist_compute_plic(mist_ctx* ctx);                   /* computePLIC!  src/VF.jl:94-103 */
/* transport!  src/Transport.jl:2-268. With solveNS = 0 the VOF part (new VF, normals, PLIC);
 * with solveNS the momentum predictor too: us, vs, ws from the cut pre-images in the band
 * (plus the 12 flux-correction tets per cell under the finite-difference scheme) and from
 * convective fluxes elsewhere, then curvature, mask, surface force, viscous fluxes, gravity
 * and the reflect boundary condition on us, vs, ws. */
int mist_transport(mist_ctx* ctx, double dt, double t);
int mist_compute_props(mist_ctx* ctx);                  /* compute_props!  src/Tools.jl:1186-1210 */
int mist_pressure_rhs(mist_ctx* ctx);                   /* RHS = div(u*)  src/Pressure.jl:16-22 */
/* Variable-density Poisson solve  A p = b,  A = dt div(rho_face^-1 grad) with wall faces
 * dropped (src/Pressure.jl:909-1033 scaling), b = RHS field, to ||b - A p||_inf <= tol.
 * Replaces pressure_solver!/mg_cycler (src/Pressure.jl:5-30, src/Multigrid.jl:333-371).
 * precision: MIST_MG_FP64 or MIST_MG_MIXED. P (MIST_FIELD_P) is the initial guess and
 * result (mean-zero gauge). */
int mist_pressure_solve(mist_ctx* ctx, double dt, double tol, int64_t precision,
                        int64_t* cycles, double* final_resid);
int mist_corrector(mist_ctx* ctx, double dt);           /* corrector!  src/Velocity.jl:3-20 */
/* Semi-Lagrangian ("SL-SL") pressure equation, pressure_scheme = semi-Lagrangian:
 *   A(P) = divg_cell of the corrected faces u* - dt/rho grad P, the SL pre-image volume in the
 *   band (src/Pressure.jl:83-114, src/Velocity.jl:72-122). mist_pressure_solve then runs the
 *   FAS cycles of mg_cycler/mg_fas! (src/Multigrid.jl:333-500) with res_iteration as smoother
 *   (diagonal-Newton steps with Anderson mixing, src/Pressure.jl:777-884) to max|A(P)| < tol,
 *   or res_iteration alone for mg_lvl = 1; precision is ignored (FP64). Non-periodic axes.
 * mist_pressure_operator: A!(P) of the current P, face velocities, densities and band into
 *   MIST_FIELD_DIVG (src/Pressure.jl:83-114).
 * mist_pressure_jacobian: jacob_single (src/Pressure.jl:699-717), the diagonal of dA/dP by
 *   central differences with delta 5, each cell's columns from the unperturbed P, into the host
 *   array out of n = the cell-extent count (interior set). */
int mist_pressure_operator(mist_ctx* ctx, double dt);
int mist_pressure_jacobian(mist_ctx* ctx, double dt, double* out, size_t n);
int mist_reduce(mist_ctx* ctx, int kind, int field, double* out); /* interior cells, FP64 */

/* One step of run_solver with solveNS=false (src/NavierStokes_Parallel.jl:125-178):
 * dt = compute_dt, t += dt, defineVelocity!(t), transport!, computeBand!.
 * t is read and advanced; dt_used returns the step. If clamp_to > 0 the step is
 * clamped so t does not pass clamp_to (test/test_deformation.jl:200-205). */
int mist_vof_step(mist_ctx* ctx, double* t, double clamp_to, double* dt_used);
/* Initialisation of run_solver for the VOF fields (src/NavierStokes_Parallel.jl:73-98):
 * Neumann!(VF), update_VF_borders!, computeBand!, computeNormal!, computePLIC!. */
int mist_vof_init(mist_ctx* ctx);

/* ---- Navier-Stokes stages (run_solver with solveNS) ------------------------------- */
/* Sides of the boundary-condition descriptor. */
enum { MIST_SIDE_XM = 0, MIST_SIDE_XP = 1, MIST_SIDE_YM = 2, MIST_SIDE_YP = 3, MIST_SIDE_ZM = 4, MIST_SIDE_ZP = 5 };
/* Wall velocity of one side (default 0). Replaces the BC! closure of the examples
 * (examples/example_viscous_bubble3D.jl:96-146, examples/example1.jl:94-136): on every
 * non-periodic side the first ghost layer of each velocity component becomes
 * -interior + 2 U_wall, axes in the order x, y, z. */
int mist_set_wall_velocity(mist_ctx* ctx, int side, const double vel[3]);
/* BC! on (MIST_FIELD_U, _V, _W) or (MIST_FIELD_US, _VS, _WS). */
int mist_apply_bc(mist_ctx* ctx, int fu, int fv, int fw);
/* interpolateFace!  src/Velocity.jl:23-37: uf, vf, wf from (u, v, w) or (us, vs, ws). */
int mist_interpolate_face(mist_ctx* ctx, int fu, int fv, int fw);
int mist_interpolate_center(mist_ctx* ctx);             /* interpolateCenter!  src/Velocity.jl:39-57 */
int mist_divergence(mist_ctx* ctx, double dt);          /* divergence!  src/Velocity.jl:59-122 */
int mist_compute_curvature(mist_ctx* ctx);              /* compute_curvature!  src/ELVIRA.jl:143-192 */
/* mask_maker! + compute_sf!  src/Tools.jl:1582-1598, src/ELVIRA.jl:256-299 (the mask is
 * never cleared, as in the reference). */
int mist_surface_force(mist_ctx* ctx);
/* Initialisation of run_solver with solveNS (src/NavierStokes_Parallel.jl:73-105):
 * Neumann!(VF), BC!(u,v,w), update_borders! of u, v, w and VF, interpolateFace!(u,v,w),
 * compute_props!, computeBand!, computeNormal!, computePLIC!, divergence! at the first dt. */
int mist_ns_init(mist_ctx* ctx);
/* The restart branch of run_solver's initialisation (src/NavierStokes_Parallel.jl:59-67 and
 * :87-100), for a context whose VF, uf, vf, wf and P were uploaded from a restart file's arrays
 * (fillArrays, src/ReadData.jl:3-53: cell extents with their ghosts) into an otherwise fresh
 * context: VF beyond the cell extents is cleared (as fillArrays leaves it), u, v, w are rebuilt
 * from the faces by interpolateCenter! (ghosts from us, vs, ws, zero in a fresh context), then
 * update_borders!, compute_props!, computeBand!, computeNormal!, computePLIC!, compute_dt and
 * divergence!. *dt (may be NULL) returns compute_dt's step. The time loop then continues with
 * mist_vof_step / mist_ns_step from the restart time. */
int mist_restart_init(mist_ctx* ctx, double* dt);
/* One step of run_solver with solveNS (src/NavierStokes_Parallel.jl:125-168): dt, transport!,
 * computeBand!, compute_props!, interpolateFace!(us,vs,ws), the pressure solve (as
 * mist_pressure_solve: b = div(u*) under the finite-difference scheme, the SL equation under
 * the semi-Lagrangian one), corrector!, interpolateCenter!, update_borders!, and divergence!. */
int mist_ns_step(mist_ctx* ctx, double* t, double clamp_to, double* dt_used, double tol, int64_t precision,
                 int64_t* cycles, double* final_resid);

/* ---- adjoint of the SL-VOF step (no reference counterpart) ------------------------- */
/* Reverse mode of one transport! step of the VOF part (solveNS = 0, 5 tets) followed by
 * computeBand!, with the band held fixed. On entry the context holds the step's input
 * state (VF with its ghosts, as the forward step found it; normals, PLIC constants and band
 * are recomputed for it when stale) and MIST_FIELD_VOF_BAR holds dL/dVF of the step's
 * output. On return MIST_FIELD_VOF_BAR holds dL/dVF of the input, and dL/d(uf, vf, wf) of
 * this step has been added to MIST_FIELD_UF_BAR, _VFACE_BAR, _WF_BAR (zero them before the
 * first step of a sweep). Branches (cut cases, clamps, ELVIRA's choice, cell searches) are
 * those of the forward evaluation. The first call allocates the work arrays. */
int mist_transport_vjp(mist_ctx* ctx, double dt);
/* Deferred overflow checks (no reference counterpart; for loops of many stage calls, such as
 * the checkpointed reverse sweeps). mist_transport and mist_transport_vjp normally wait for the
 * device after each call to report an overflow of the cut walk (MIST_ERR_OVERFLOW). With
 * on != 0 they return without waiting; the overflow is then reported by the next call that
 * waits for the device anyway (mist_download, mist_reduce, mist_compute_dt, mist_ns_step, a
 * mist_vof_step that follows another call) and by mist_set_deferred_checks(ctx, 0), which
 * waits and returns it. Consecutive mist_vof_step calls with no other call in between do not
 * wait for the previous step either: each reduces the velocity maxima right behind its
 * defineVelocity! and the next one waits only for those (the same dt bit for bit). */
int mist_set_deferred_checks(mist_ctx* ctx, int on);

/* Step checkpoints: copy the state a step starts from (VF with its ghosts, the band flags
 * and band-cell list, the normals, PLIC constants and ELVIRA's choice; those of them that are
 * stale are computed first) to device slot `slot` (0 <= slot < 4096, allocated on first use,
 * about 5 cell fields per slot) and back. */
int mist_checkpoint_save(mist_ctx* ctx, int slot);
int mist_checkpoint_restore(mist_ctx* ctx, int slot);

/* computeDist of a batch, as mist_plic_dist_batch, without allocating: no hipMalloc, hipFree
 * or allocating library call. No reference counterpart; it is for callers that cannot
 * allocate device memory (tight serving loops, code that forbids allocation). The batch is
 * staged in a per-GPU device array of 4 MiB, held until the process ends, so m may be at most
 * 65536 (larger batches fail with MIST_ERR_PARAM). Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (from several host threads) race. */
int mist_plic_dist_batch_noalloc(mist_ctx* ctx, const int64_t* ijk, const double* nrm,
                                 const double* vf, double* dist_out, int64_t m);

/* ELVIRA's choice for the current VF (ELVIRA!'s argmin, src/ELVIRA.jl:106-137), for checks
 * against another implementation: out[(i-1) + Nx*((j-1) + Ny*(k-1))] = the candidate kept in
 * interface cell (i, j, k), numbered c = 9*family + 3*outer pair + inner pair in the loop
 * order of ELVIRA_calc! (families X-Y, X-Z, Y-Z; src/ELVIRA.jl:71-98), -1 in other cells.
 * n must be Nx*Ny*Nz. No reference counterpart: the reference keeps the choice implicit in
 * the normals. Synchronises. */
int mist_elvira_choice(mist_ctx* ctx, int32_t* out, size_t n);

/* Batched PLIC primitives (first slice; host arrays of length m).
 * computeDist (src/VF.jl:109-156) for cells (i,j,k) with normals n and fractions vf. */
int mist_plic_dist_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm,
                         const double* vf, double* dist_out, int64_t m);
/* computePLIC2VF (src/VF.jl:219-278): liquid fraction of cell (i,j,k) under plane (n, D). */
int mist_plic_vf_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm,
                       const double* dist, double* vf_out, int64_t m);

/* Per-stage device time of the last calls (ms), for profiling: kinds are the stage
 * entry points in order normal, transport, plic, band, mg. Filled only when the process
 * runs with MIST_STAGE_TIMING set (each stage then synchronises). */
int mist_stage_times(const mist_ctx* ctx, double* ms, int n);

/* Kernel profile: HIP events on the context's stream around every launch of the kernels
 * below while enabled, and device-side totals of the work units each launch processed.
 * No reference counterpart (the reference has only @time around whole runs); used by the
 * benchmark to compute achieved FLOP/s and bytes/s per kernel. */
enum { MIST_PROF_ELVIRA = 0,     /* units: interface cells scored */
       MIST_PROF_SL_CUT = 1,     /* units: band cells transported */
       MIST_PROF_PLIC = 2,       /* units: cells (incl. ghosts) */
       MIST_PROF_PROJECT = 3,    /* units: cell vertices visited */
       MIST_PROF_MG_RBGS = 4,    /* units: cells of the finest level per red-black sweep */
       MIST_PROF_MG_RESID = 5,   /* units: cells */
       MIST_PROF_SL_GENERAL = 6, /* units: band cells left to the general cut walk */
       MIST_PROF_NS_CUT = 7,     /* units: band cells of the momentum predictor */
       MIST_PROF_NS_FIELDS = 8,  /* units: interior cells per stencil pass of the predictor */
       MIST_PROF_COUNT = 9 };
int mist_profile_enable(mist_ctx* ctx, int on);   /* on != 0 also resets the totals */
int mist_profile_read(mist_ctx* ctx, int kernel, double* total_ms, int64_t* launches, int64_t* units);

#ifdef __cplusplus
}
#endif

#endif /* MIST_HIP_H */
