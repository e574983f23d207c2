// mist_capi.hip -- the C ABI of libmist_hip.so (declared in include/mist_hip.h): context
// life cycle, field transfer, and one entry point per reference stage.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mist_internal.h"

static std::string g_last_error;

// Device-resident staging of mist_plic_dist_batch_noalloc: one array per GPU, held for the
// life of the process. Per batch entry 64 bytes (3 int64 cell indices, 3 normal components,
// VF, the result), so 4 MiB hold NOALLOC_BATCH_MAX entries. Host code must never pass the
// array's name to a kernel or copy: its device address comes from hipGetSymbolAddress, taken
// once per device in mist_ctx_create and kept in g_scratch_ptr.
namespace {
constexpr int64_t NOALLOC_BATCH_MAX = 65536;
constexpr size_t SCRATCH_DOUBLES = (size_t)NOALLOC_BATCH_MAX * 8;
constexpr int MAX_DEVICES = 64;
void* g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu;
}  // namespace
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// The device address of this device's scratch (the current device must be `device`).
static hipError_t scratch_address(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = ptr;
    return e;
}

static int fail(mist_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    g_last_error = msg;
    return code;
}

// VF changed: the cached normals, PLIC constants and band list of the old VF no longer apply
// (the version counters make the normals and constants stale; the band list is flagged).
static inline void vf_changed(mist_ctx* ctx) {
    ctx->vf_version++;
    ctx->band_list_valid = false;
}

// What a write to `field` (upload, or through the device pointer) invalidates.
static void field_written(mist_ctx* ctx, int field) {
    if (field == MIST_FIELD_VOF) vf_changed(ctx);
    if (field == MIST_FIELD_NX || field == MIST_FIELD_NY || field == MIST_FIELD_NZ) ctx->normals_version = 0;
    if (field == MIST_FIELD_NX || field == MIST_FIELD_NY || field == MIST_FIELD_NZ || field == MIST_FIELD_PLIC_D)
        ctx->plic_version = 0;
    if (field == MIST_FIELD_BAND) ctx->band_list_valid = false;
}

#define CHECK_CTX()                                                   \
    do {                                                              \
        if (!ctx) return fail(nullptr, MIST_ERR_PARAM, "null context"); \
        ++ctx->api_seq;                                               \
    } while (0)

#define RET_IF(x)                                 \
    do {                                          \
        int st_ = (x);                            \
        if (st_) { g_last_error = ctx->err; return st_; } \
    } while (0)

namespace {

struct StageTimer {
    mist_ctx* ctx;
    int stage;
    StageTimer(mist_ctx* c, int s) : ctx(c), stage(s) {
        if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
    }
    ~StageTimer() {
        if (ctx->timing) {
            (void)hipEventRecord(ctx->ev1, ctx->stream);
            (void)hipEventSynchronize(ctx->ev1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
            ctx->stage_ms[stage] += ms;
        }
    }
};

int check_params(const mist_params* p, std::string& err) {
    if (p->Nx < 1 || p->Ny < 1 || p->Nz < 1) { err = "Nx, Ny, Nz must be >= 1"; return MIST_ERR_PARAM; }
    if ((int64_t)p->Nx * p->Ny * p->Nz > (int64_t)1 << 30) { err = "grid too large for 32-bit cell lists"; return MIST_ERR_PARAM; }
    if (!(p->Lx > 0 && p->Ly > 0 && p->Lz > 0)) { err = "Lx, Ly, Lz must be > 0"; return MIST_ERR_PARAM; }
    if (p->tesselation != MIST_TESS_5_TETS) { err = "incompatible tesselation called"; return MIST_ERR_PARAM; }
    if (p->normal_method != MIST_NORMAL_ELVIRA) { err = "Unknown method to copute interface normal"; return MIST_ERR_PARAM; }
    if (p->projection_method < 0 || p->projection_method > 2) { err = "Unknown projection_method in project!"; return MIST_ERR_PARAM; }
    if (!(p->CFL > 0)) { err = "CFL must be > 0"; return MIST_ERR_PARAM; }
    if (p->mg_coarse_op != MIST_COARSE_FACE_AVERAGE && p->mg_coarse_op != MIST_COARSE_RESTRICTED_VF) { err = "unknown mg_coarse_op"; return MIST_ERR_PARAM; }
    if (p->mg_nu1 < 0 || p->mg_nu2 < 0 || p->mg_coarse_sweeps < 0) { err = "multigrid sweep counts must be >= 0"; return MIST_ERR_PARAM; }
    return MIST_OK;
}

// create_mesh coordinates (src/Mesh.jl:76-108): range(0, L, N+1) then ghosts and centres.
void axis_coords(int64_t N, double L, int64_t ng, std::vector<double>& fc, std::vector<double>& cc, double& h) {
    fc.assign((size_t)(N + 2 * ng + 1), 0.0);
    for (int64_t q = 0; q <= N; ++q) fc[(size_t)(ng + q)] = ((double)q * L) / (double)N;
    h = fc[(size_t)(ng + 1)] - fc[(size_t)ng];
    for (int64_t n = 1; n <= ng; ++n) {
        fc[(size_t)(ng - n)] = fc[(size_t)ng] - (double)n * h;
        fc[(size_t)(ng + N + n)] = fc[(size_t)(ng + N)] + (double)n * h;
    }
    cc.assign((size_t)(N + 2 * ng), 0.0);
    for (size_t q = 0; q < cc.size(); ++q) cc[q] = 0.5 * (fc[q] + fc[q + 1]);
}

}  // namespace

namespace mist {

__global__ void k_accum_units(unsigned long long* acc, const int* units) { *acc += (unsigned long long)*units; }
__global__ void k_zero_u64(unsigned long long* a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0ull;
}

static hipEvent_t prof_event(mist_ctx* ctx) {
    KernelProfile& P = ctx->prof;
    if (!P.pool.empty()) {
        hipEvent_t e = P.pool.back();
        P.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

hipEvent_t prof_begin(mist_ctx* ctx, int kernel) {
    if (!ctx->prof.on) return nullptr;
    hipEvent_t a = prof_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
    return a;
}

void prof_end(mist_ctx* ctx, int kernel, hipEvent_t a, const int* device_units, int64_t host_units) {
    KernelProfile& P = ctx->prof;
    if (!P.on || !a) return;
    hipEvent_t b = prof_event(ctx);
    if (!b) return;
    (void)hipEventRecord(b, ctx->stream);
    P.pending[kernel].push_back({a, b});
    P.launches[kernel] += 1;
    P.host_units[kernel] += host_units;
    if (device_units && P.d_units)
        hipLaunchKernelGGL(k_accum_units, dim3(1), dim3(1), 0, ctx->stream, P.d_units + kernel, device_units);
}

}  // namespace mist

extern "C" {

int mist_params_default(mist_params* p) {
    if (!p) return MIST_ERR_PARAM;
    std::memset(p, 0, sizeof(*p));
    p->max_dt = INFINITY;
    p->solveNS = 1;
    p->pressure_scheme = MIST_PS_SEMI_LAGRANGIAN;
    p->mg_lvl = 1;
    p->VFlo = 1e-10;
    p->VFhi = 1.0 - 1e-10;
    p->nband = -1;
    p->vel_period = 2.0;
    p->mg_precision = MIST_MG_FP64;
    p->mg_nu1 = 10;
    p->mg_nu2 = 10;
    p->mg_max_cycles = 10000;
    p->mg_coarse_op = MIST_COARSE_FACE_AVERAGE;
    p->mg_coarse_sweeps = 64;
    return MIST_OK;
}

const char* mist_last_error(const mist_ctx* ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }

int mist_ctx_destroy(mist_ctx* ctx) {
    if (!ctx) return MIST_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    mist::mg_free(ctx->mg);
    mist::fas_free(ctx->fas);
    if (ctx->vjp_exec) (void)hipGraphExecDestroy(ctx->vjp_exec);
    if (ctx->vjp_graph) (void)hipGraphDestroy(ctx->vjp_graph);
    for (int f = 0; f < MIST_FIELD_COUNT; ++f)
        if (ctx->f[f]) (void)hipFree(ctx->f[f]);
    void* bufs[] = {ctx->geom, ctx->trig, ctx->pv, ctx->band_list, ctx->iface_list, ctx->defer_list, ctx->cut_list,
                    ctx->counters, ctx->red, ctx->ns_sums, ctx->elv_best, ctx->adj_lam, ctx->adj_plane, ctx->adj_pv,
                    ctx->adj_sums, ctx->elv_dist, ctx->vtx_list, ctx->elv_ties};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (double* b : ctx->ckpt)
        if (b) (void)hipFree(b);
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    if (ctx->h_red) (void)hipHostFree(ctx->h_red);
    if (ctx->h_vmax) (void)hipHostFree(ctx->h_vmax);
    if (ctx->vmax_ev) (void)hipEventDestroy(ctx->vmax_ev);
    for (int k = 0; k < MIST_PROF_COUNT; ++k)
        for (auto& pr : ctx->prof.pending[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (hipEvent_t e : ctx->prof.pool) (void)hipEventDestroy(e);
    if (ctx->prof.d_units) (void)hipFree(ctx->prof.d_units);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MIST_OK;
}

int mist_ctx_create(const mist_params* pin, int device, mist_ctx** out) {
    if (!pin || !out) return fail(nullptr, MIST_ERR_PARAM, "null argument");
    *out = nullptr;
    std::string err;
    int st = check_params(pin, err);
    if (st) return fail(nullptr, st, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        return fail(nullptr, MIST_ERR_NODEVICE, "no HIP device " + std::to_string(device));
    mist_ctx* ctx = new mist_ctx();
    ctx->p = *pin;
    if (ctx->p.nband < 0) ctx->p.nband = (int64_t)std::ceil(ctx->p.CFL);
    ctx->device = device;
    ctx->timing = std::getenv("MIST_STAGE_TIMING") != nullptr;
    auto bail = [&](int code, const std::string& msg) {
        g_last_error = msg;
        mist_ctx_destroy(ctx);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(MIST_ERR_HIP, "hipSetDevice failed");
    {
        const hipError_t e = scratch_address(device);
        if (e != hipSuccess) return bail(MIST_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return bail(MIST_ERR_HIP, "stream");
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) return bail(MIST_ERR_HIP, "event");

    const mist_params& p = ctx->p;
    const int64_t ng = 3 + std::max<int64_t>(0, (int64_t)std::ceil(p.CFL) - 2);
    double dx, dy, dz;
    axis_coords(p.Nx, p.Lx, ng, ctx->hx, ctx->hxm, dx);
    axis_coords(p.Ny, p.Ly, ng, ctx->hy, ctx->hym, dy);
    axis_coords(p.Nz, p.Lz, ng, ctx->hz, ctx->hzm, dz);
    std::vector<double> g;
    const size_t off[6] = {0, ctx->hx.size(), ctx->hx.size() + ctx->hy.size(),
                           ctx->hx.size() + ctx->hy.size() + ctx->hz.size(), 0, 0};
    g.insert(g.end(), ctx->hx.begin(), ctx->hx.end());
    g.insert(g.end(), ctx->hy.begin(), ctx->hy.end());
    g.insert(g.end(), ctx->hz.begin(), ctx->hz.end());
    const size_t offm = g.size();
    g.insert(g.end(), ctx->hxm.begin(), ctx->hxm.end());
    g.insert(g.end(), ctx->hym.begin(), ctx->hym.end());
    g.insert(g.end(), ctx->hzm.begin(), ctx->hzm.end());
    if (hipMalloc(&ctx->geom, g.size() * sizeof(double)) != hipSuccess) return bail(MIST_ERR_OOM, "geometry");
    if (hipMemcpy(ctx->geom, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bail(MIST_ERR_HIP, "geometry copy");
    mist::DevMesh& m = ctx->m;
    m.Nx = (int)p.Nx; m.Ny = (int)p.Ny; m.Nz = (int)p.Nz; m.ng = (int)ng;
    m.cnx = (int)(p.Nx + 2 * ng); m.cny = (int)(p.Ny + 2 * ng); m.cnz = (int)(p.Nz + 2 * ng);
    m.vnx = m.cnx + 6; m.vny = m.cny + 6; m.vnz = m.cnz + 6;
    m.dx = dx; m.dy = dy; m.dz = dz;
    m.Lx = p.Lx; m.Ly = p.Ly; m.Lz = p.Lz;
    const int64_t imino = 1 - ng;  // pointers pre-offset so that x[i] is reference index i
    m.x = ctx->geom + off[0] - imino;
    m.y = ctx->geom + off[1] - imino;
    m.z = ctx->geom + off[2] - imino;
    m.xm = ctx->geom + offm - imino;
    m.ym = ctx->geom + offm + ctx->hxm.size() - imino;
    m.zm = ctx->geom + offm + ctx->hxm.size() + ctx->hym.size() - imino;
    {
        const int64_t G = (int64_t)g.size();
        if (hipMalloc(&ctx->trig, 3 * G * sizeof(double)) != hipSuccess) return bail(MIST_ERR_OOM, "trig tables");
        if (mist::launch_trig_tables(ctx, G)) return bail(MIST_ERR_HIP, "trig tables");
        const size_t offa[6] = {off[0], off[1], off[2], offm, offm + ctx->hxm.size(),
                                offm + ctx->hxm.size() + ctx->hym.size()};
        for (int a = 0; a < 6; ++a) {
            ctx->vt.s1[a] = ctx->trig + offa[a] - imino;
            ctx->vt.s2[a] = ctx->trig + G + offa[a] - imino;
            ctx->vt.c1[a] = ctx->trig + 2 * G + offa[a] - imino;
        }
    }
    mist::DevParams& dp = ctx->dp;
    dp.VFlo = p.VFlo; dp.VFhi = p.VFhi; dp.nband = (int)p.nband;
    dp.projection = (int)p.projection_method; dp.zalesak = p.test_case == MIST_CASE_ZALESAK;
    dp.rho_liq = p.rho_liq; dp.rho_gas = p.rho_gas;

    // fields
    for (int f = 0; f < MIST_FIELD_COUNT; ++f) {
        int64_t lo = imino, n[3] = {m.cnx, m.cny, m.cnz}, el = 8;
        if (f == MIST_FIELD_VOF || f == MIST_FIELD_VOF_BAR) { lo = imino - 3; n[0] += 6; n[1] += 6; n[2] += 6; }
        if (f == MIST_FIELD_DENX || f == MIST_FIELD_VISCX || f == MIST_FIELD_SFX) n[0] += 1;
        if (f == MIST_FIELD_DENY || f == MIST_FIELD_VISCY || f == MIST_FIELD_SFY) n[1] += 1;
        if (f == MIST_FIELD_DENZ || f == MIST_FIELD_VISCZ || f == MIST_FIELD_SFZ) n[2] += 1;
        if (f == MIST_FIELD_BAND || f == MIST_FIELD_MASK) el = 2;
        for (int d = 0; d < 3; ++d) { ctx->flo[f][d] = lo; ctx->fn[f][d] = n[d]; }
        ctx->felem[f] = el;
        const size_t bytes = (size_t)(n[0] * n[1] * n[2] * el);
        if (hipMalloc(&ctx->f[f], bytes) != hipSuccess) return bail(MIST_ERR_OOM, "field allocation");
        if (hipMemset(ctx->f[f], 0, bytes) != hipSuccess) return bail(MIST_ERR_HIP, "field clear");
    }
    const int64_t ncell = p.Nx * p.Ny * p.Nz;
    const int64_t nvtx = (p.Nx + 1) * (p.Ny + 1) * (p.Nz + 1);
    ctx->red_blocks = 1024;
    if (hipMalloc(&ctx->pv, (size_t)nvtx * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->vtx_list, (size_t)nvtx * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->band_list, (size_t)ncell * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->iface_list, (size_t)ncell * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->defer_list, (size_t)ncell * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->cut_list, (size_t)ncell * 24) != hipSuccess ||
        hipMalloc(&ctx->elv_best, (size_t)ncell * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->counters, 16 * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->red, (size_t)(4 * ctx->red_blocks + 16) * sizeof(double)) != hipSuccess)
        return bail(MIST_ERR_OOM, "work allocation");
    if (p.solveNS && hipMalloc(&ctx->ns_sums, (size_t)ncell * 6 * sizeof(double)) != hipSuccess)
        return bail(MIST_ERR_OOM, "work allocation");
    if (hipMemset(ctx->counters, 0, 16 * sizeof(int)) != hipSuccess) return bail(MIST_ERR_HIP, "clear");
    if (hipHostMalloc(&ctx->h_counters, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&ctx->h_red, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&ctx->h_vmax, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(MIST_ERR_OOM, "pinned allocation");
    if (hipEventCreateWithFlags(&ctx->vmax_ev, hipEventDisableTiming) != hipSuccess)
        return bail(MIST_ERR_HIP, "event");
    *out = ctx;
    return MIST_OK;
}

int mist_field_extents(const mist_ctx* ctx, int field, int64_t lo[3], int64_t n[3], int64_t* elem) {
    if (!ctx) return MIST_ERR_PARAM;
    if (field < 0 || field >= MIST_FIELD_COUNT) return MIST_ERR_PARAM;
    for (int d = 0; d < 3; ++d) { lo[d] = ctx->flo[field][d]; n[d] = ctx->fn[field][d]; }
    if (elem) *elem = ctx->felem[field];
    return MIST_OK;
}

static size_t field_bytes(const mist_ctx* ctx, int f) {
    return (size_t)(ctx->fn[f][0] * ctx->fn[f][1] * ctx->fn[f][2] * ctx->felem[f]);
}

int mist_field_device_ptr(mist_ctx* ctx, int field, void** dptr) {
    CHECK_CTX();
    if (field < 0 || field >= MIST_FIELD_COUNT || !dptr) return fail(ctx, MIST_ERR_PARAM, "bad field");
    *dptr = ctx->f[field];
    // the caller may write through the pointer: cached derived state is no longer trusted
    field_written(ctx, field);
    return MIST_OK;
}

int mist_field_modified(mist_ctx* ctx, int field) {
    CHECK_CTX();
    if (field < 0 || field >= MIST_FIELD_COUNT) return fail(ctx, MIST_ERR_PARAM, "bad field");
    field_written(ctx, field);
    return MIST_OK;
}

int mist_upload(mist_ctx* ctx, int field, const void* host, size_t nbytes) {
    CHECK_CTX();
    if (field < 0 || field >= MIST_FIELD_COUNT || !host) return fail(ctx, MIST_ERR_PARAM, "bad field");
    if (nbytes != field_bytes(ctx, field))
        return fail(ctx, MIST_ERR_PARAM, "size mismatch: expected " + std::to_string(field_bytes(ctx, field)) + " bytes");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    MIST_HIP_TRY(hipMemcpyAsync(ctx->f[field], host, nbytes, hipMemcpyHostToDevice, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    field_written(ctx, field);
    return MIST_OK;
}

static int check_overflow(mist_ctx* ctx, bool force = false);

int mist_download(mist_ctx* ctx, int field, void* host, size_t nbytes) {
    CHECK_CTX();
    if (field < 0 || field >= MIST_FIELD_COUNT || !host) return fail(ctx, MIST_ERR_PARAM, "bad field");
    if (nbytes != field_bytes(ctx, field)) return fail(ctx, MIST_ERR_PARAM, "size mismatch");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    MIST_HIP_TRY(hipMemcpyAsync(host, ctx->f[field], nbytes, hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->defer_checks) RET_IF(check_overflow(ctx, true));
    return MIST_OK;
}

int mist_set_deferred_checks(mist_ctx* ctx, int on) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    const bool was = ctx->defer_checks;
    ctx->defer_checks = on != 0;
    if (was && !on) RET_IF(check_overflow(ctx, true));  // what the deferred calls left
    return MIST_OK;
}

int mist_init_vof(mist_ctx* ctx, int shape, const double* args, int nargs) {
    CHECK_CTX();
    if (!args || shape < 0 || shape > 2 || nargs < (shape == MIST_IC_CIRCLE ? 3 : 4))
        return fail(ctx, MIST_ERR_PARAM, "bad initial-condition shape or arguments");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_init_vof(ctx, shape, args));
    vf_changed(ctx);
    return MIST_OK;
}

// face arrays (one extra face along their axis) and the int16 mask take no ghost fills
static bool face_field(int f) {
    return f == MIST_FIELD_DENX || f == MIST_FIELD_DENY || f == MIST_FIELD_DENZ || f == MIST_FIELD_VISCX ||
           f == MIST_FIELD_VISCY || f == MIST_FIELD_VISCZ || f == MIST_FIELD_SFX || f == MIST_FIELD_SFY ||
           f == MIST_FIELD_SFZ;
}

int mist_neumann(mist_ctx* ctx, int field) {
    CHECK_CTX();
    if (field < 0 || field >= MIST_FIELD_COUNT || face_field(field) || field == MIST_FIELD_MASK)
        return fail(ctx, MIST_ERR_PARAM, "Neumann! applies to cell fields");
    RET_IF(mist::launch_neumann(ctx, field));
    if (field == MIST_FIELD_VOF) vf_changed(ctx);
    return MIST_OK;
}

int mist_update_borders(mist_ctx* ctx, int field) {
    CHECK_CTX();
    if (field < 0 || field >= MIST_FIELD_COUNT || face_field(field) || field == MIST_FIELD_MASK)
        return fail(ctx, MIST_ERR_PARAM, "update_borders! applies to cell fields");
    RET_IF(mist::launch_borders(ctx, field));
    if (field == MIST_FIELD_VOF && (ctx->p.xper || ctx->p.yper || ctx->p.zper)) vf_changed(ctx);
    return MIST_OK;
}

int mist_define_velocity(mist_ctx* ctx, double t) {
    CHECK_CTX();
    if (ctx->p.vf_velocity < 0 || ctx->p.vf_velocity > MIST_VEL_AIRBLAST || ctx->p.vf_velocity == MIST_VEL_NONE)
        return fail(ctx, MIST_ERR_PARAM, "Unknown VFVelocity");
    RET_IF(mist::launch_define_velocity(ctx, t));
    return MIST_OK;
}

// compute_dt (src/Tools.jl:317-337)
static double dt_from_max(const mist_ctx* ctx, const double mx[3]) {
    using mist::EPS;
    auto jmin = [](double a, double b) {
        if (a != a) return a;
        if (b != b) return b;
        if (a == b) return std::signbit(a) ? a : b;
        return a < b ? a : b;
    };
    auto jmax = [](double a, double b) {
        if (a != a) return a;
        if (b != b) return b;
        if (a == b) return std::signbit(a) ? b : a;
        return a > b ? a : b;
    };
    const mist::DevMesh& m = ctx->m;
    const double convec = jmin(jmin(m.dx / mx[0], m.dy / mx[1]), m.dz / mx[2]);
    const double viscous = jmin(jmin(m.dx, m.dy), m.dz) / jmax(ctx->p.mu_liq, ctx->p.mu_gas);
    return jmin(ctx->p.max_dt, ctx->p.CFL * jmin(convec, viscous));
}

int mist_compute_dt(mist_ctx* ctx, double* dt) {
    CHECK_CTX();
    if (!dt) return fail(ctx, MIST_ERR_PARAM, "null dt");
    double mx[3];
    RET_IF(mist::launch_maxabs_uvw(ctx, mx));
    if (ctx->defer_checks) RET_IF(check_overflow(ctx, true));
    *dt = dt_from_max(ctx, mx);
    return MIST_OK;
}

int mist_compute_band(mist_ctx* ctx) {
    CHECK_CTX();
    StageTimer tm(ctx, MIST_STAGE_BAND);
    RET_IF(mist::launch_band(ctx));
    ctx->band_list_valid = true;
    return MIST_OK;
}

int mist_compute_normal(mist_ctx* ctx) {
    CHECK_CTX();
    // ELVIRA is a pure function of VF: normals computed for this VF are reused
    if (ctx->normals_version == ctx->vf_version) return MIST_OK;
    StageTimer tm(ctx, MIST_STAGE_NORMAL);
    RET_IF(mist::launch_normals(ctx));
    ctx->normals_version = ctx->vf_version;
    return MIST_OK;
}

int mist_compute_plic(mist_ctx* ctx) {
    CHECK_CTX();
    StageTimer tm(ctx, MIST_STAGE_PLIC);
    RET_IF(mist::launch_plic(ctx));
    // D is this VF's when the normals are (they are recomputed whenever VF changes)
    ctx->plic_version = ctx->normals_version == ctx->vf_version ? ctx->vf_version : 0;
    return MIST_OK;
}

// The cut walk's overflow flag (sticky) read back on the host; with deferred checks, only where
// the call waits for the device anyway (`force`).
static int check_overflow(mist_ctx* ctx, bool force) {
    if (ctx->defer_checks && !force) return MIST_OK;
    MIST_HIP_TRY(hipMemcpyAsync(ctx->h_counters, ctx->counters, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->h_counters[2]) {
        ctx->err = "cutTet: pre-image walk left the mesh or needed more than 16 cut levels";
        return MIST_ERR_OVERFLOW;
    }
    return MIST_OK;
}

static int transport_body(mist_ctx* ctx, double dt) {
    RET_IF(mist_compute_normal(ctx));
    if (!ctx->band_list_valid) RET_IF(mist_compute_band(ctx));
    {
        StageTimer tm(ctx, MIST_STAGE_TRANSPORT);
        RET_IF(mist::launch_transport(ctx, dt));
    }
    vf_changed(ctx);
    RET_IF(mist_compute_normal(ctx));
    RET_IF(mist_compute_plic(ctx));
    ctx->band_list_valid = false;  // band no longer matches VF; the caller recomputes it
    return MIST_OK;
}

// transport! with the momentum predictor (src/Transport.jl:2-268)
static int transport_ns_body(mist_ctx* ctx, double dt) {
    RET_IF(mist_compute_normal(ctx));
    if (!ctx->band_list_valid) RET_IF(mist_compute_band(ctx));
    {
        StageTimer tm(ctx, MIST_STAGE_TRANSPORT);
        RET_IF(mist::launch_ns_advect(ctx, dt));
    }
    vf_changed(ctx);
    RET_IF(mist_compute_normal(ctx));
    {
        StageTimer tm(ctx, MIST_STAGE_TRANSPORT);
        RET_IF(mist::launch_ns_forces(ctx, dt));
        RET_IF(mist::launch_reflect_bc(ctx, MIST_FIELD_US, MIST_FIELD_VS, MIST_FIELD_WS));
    }
    RET_IF(mist_compute_plic(ctx));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_US));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_VS));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_WS));
    ctx->band_list_valid = false;
    return MIST_OK;
}

int mist_transport(mist_ctx* ctx, double dt, double t) {
    CHECK_CTX();
    (void)t;
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(ctx->p.solveNS ? transport_ns_body(ctx, dt) : transport_body(ctx, dt));
    RET_IF(check_overflow(ctx));
    return MIST_OK;
}

int mist_vof_init(mist_ctx* ctx) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_neumann(ctx, MIST_FIELD_VOF));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_VOF));
    vf_changed(ctx);
    RET_IF(mist_compute_band(ctx));
    RET_IF(mist_compute_normal(ctx));
    RET_IF(mist_compute_plic(ctx));
    return MIST_OK;
}

int mist_vof_step(mist_ctx* ctx, double* t, double clamp_to, double* dt_used) {
    CHECK_CTX();
    if (!t) return fail(ctx, MIST_ERR_PARAM, "null t");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    double mx[3];
    // with deferred checks, the maxima the previous step reduced behind its velocity (when no
    // other call came in between): wait for them, not for the previous step's transport
    const bool pending = ctx->defer_checks && ctx->vmax_pending && ctx->api_seq == ctx->vmax_seq + 1;
    ctx->vmax_pending = false;
    if (pending) {
        MIST_HIP_TRY(hipEventSynchronize(ctx->vmax_ev));
        mx[0] = ctx->h_vmax[0]; mx[1] = ctx->h_vmax[1]; mx[2] = ctx->h_vmax[2];
    } else {
        RET_IF(mist::launch_maxabs_uvw(ctx, mx));  // synchronises: the previous step is complete
        RET_IF(check_overflow(ctx, true));
    }
    double dt = dt_from_max(ctx, mx);
    if (clamp_to > 0.0 && *t + dt > clamp_to) dt = clamp_to - *t;
    *t += dt;
    if (ctx->defer_checks) RET_IF(mist::launch_define_velocity_vmax(ctx, *t));
    else RET_IF(mist::launch_define_velocity(ctx, *t));
    RET_IF(transport_body(ctx, dt));
    RET_IF(mist_compute_band(ctx));
    if (ctx->defer_checks) {
        ctx->vmax_pending = true;
        ctx->vmax_seq = ctx->api_seq;
    }
    if (dt_used) *dt_used = dt;
    return MIST_OK;
}

int mist_compute_props(mist_ctx* ctx) {
    CHECK_CTX();
    RET_IF(mist::launch_props(ctx));
    return MIST_OK;
}

int mist_elvira_choice(mist_ctx* ctx, int32_t* out, size_t n) {
    CHECK_CTX();
    const mist::DevMesh& m = ctx->m;
    const size_t ncell = (size_t)m.Nx * m.Ny * m.Nz;
    if (!out || n != ncell) return fail(ctx, MIST_ERR_PARAM, "elvira choice: expected Nx*Ny*Nz entries");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist_compute_normal(ctx));  // the choice of the current VF
    std::vector<int> best(ncell);
    std::vector<double> vf((size_t)m.vnx * m.vny * m.vnz);
    MIST_HIP_TRY(hipMemcpyAsync(best.data(), ctx->elv_best, ncell * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(vf.data(), ctx->f[MIST_FIELD_VOF], vf.size() * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int k = 1; k <= m.Nz; ++k)
        for (int j = 1; j <= m.Ny; ++j)
            for (int i = 1; i <= m.Nx; ++i) {
                const size_t q = (size_t)(i - 1) + (size_t)m.Nx * ((j - 1) + (size_t)m.Ny * (k - 1));
                const double v = vf[(size_t)m.vof(i, j, k)];
                // the kernel writes only interface cells; the low byte is the candidate
                out[q] = (v >= ctx->p.VFlo && v <= ctx->p.VFhi && best[q] >= 0) ? (best[q] & 255) : -1;
            }
    return MIST_OK;
}

// ---- Navier-Stokes stages ------------------------------------------------------------
int mist_set_wall_velocity(mist_ctx* ctx, int side, const double vel[3]) {
    CHECK_CTX();
    if (side < 0 || side > 5 || !vel) return fail(ctx, MIST_ERR_PARAM, "side must be 0..5 and vel non-null");
    for (int q = 0; q < 3; ++q) ctx->bc_wall[side][q] = vel[q];
    return MIST_OK;
}

static bool velocity_triple(int fu, int fv, int fw) {
    return (fu == MIST_FIELD_U && fv == MIST_FIELD_V && fw == MIST_FIELD_W) ||
           (fu == MIST_FIELD_US && fv == MIST_FIELD_VS && fw == MIST_FIELD_WS);
}

int mist_apply_bc(mist_ctx* ctx, int fu, int fv, int fw) {
    CHECK_CTX();
    if (!velocity_triple(fu, fv, fw)) return fail(ctx, MIST_ERR_PARAM, "BC! takes (u, v, w) or (us, vs, ws)");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_reflect_bc(ctx, fu, fv, fw));
    return MIST_OK;
}

int mist_interpolate_face(mist_ctx* ctx, int fu, int fv, int fw) {
    CHECK_CTX();
    if (!velocity_triple(fu, fv, fw))
        return fail(ctx, MIST_ERR_PARAM, "interpolateFace! takes (u, v, w) or (us, vs, ws)");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_interp_face(ctx, fu, fv, fw));
    return MIST_OK;
}

int mist_interpolate_center(mist_ctx* ctx) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_interp_center(ctx));
    return MIST_OK;
}

int mist_divergence(mist_ctx* ctx, double dt) {
    CHECK_CTX();
    if (!(dt > 0.0)) return fail(ctx, MIST_ERR_PARAM, "dt must be > 0");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN && !ctx->band_list_valid) RET_IF(mist_compute_band(ctx));
    RET_IF(mist::launch_divergence(ctx, dt));
    return MIST_OK;
}

int mist_compute_curvature(mist_ctx* ctx) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_curvature(ctx));
    return MIST_OK;
}

int mist_surface_force(mist_ctx* ctx) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_surface_force(ctx));
    return MIST_OK;
}

int mist_ns_init(mist_ctx* ctx) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::launch_neumann(ctx, MIST_FIELD_VOF));
    RET_IF(mist::launch_reflect_bc(ctx, MIST_FIELD_U, MIST_FIELD_V, MIST_FIELD_W));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_U));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_V));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_W));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_VOF));
    vf_changed(ctx);
    RET_IF(mist::launch_interp_face(ctx, MIST_FIELD_U, MIST_FIELD_V, MIST_FIELD_W));
    RET_IF(mist::launch_props(ctx));
    RET_IF(mist_compute_band(ctx));
    RET_IF(mist_compute_normal(ctx));
    RET_IF(mist_compute_plic(ctx));
    double mx[3];
    RET_IF(mist::launch_maxabs_uvw(ctx, mx));
    RET_IF(mist::launch_divergence(ctx, dt_from_max(ctx, mx)));
    return MIST_OK;
}

int mist_restart_init(mist_ctx* ctx, double* dt) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    // fillArrays fills the cell extents; the VF layers beyond them stay as initArrays left them
    RET_IF(mist::launch_clear_vof_outer(ctx));
    vf_changed(ctx);
    RET_IF(mist::launch_interp_center(ctx));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_U));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_V));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_W));
    RET_IF(mist::launch_props(ctx));
    RET_IF(mist_compute_band(ctx));
    RET_IF(mist_compute_normal(ctx));
    RET_IF(mist_compute_plic(ctx));
    double mx[3];
    RET_IF(mist::launch_maxabs_uvw(ctx, mx));
    const double d = dt_from_max(ctx, mx);
    RET_IF(mist::launch_divergence(ctx, d));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (dt) *dt = d;
    return MIST_OK;
}

int mist_ns_step(mist_ctx* ctx, double* t, double clamp_to, double* dt_used, double tol, int64_t precision,
                 int64_t* cycles, double* final_resid) {
    CHECK_CTX();
    if (!t) return fail(ctx, MIST_ERR_PARAM, "null t");
    if (precision != MIST_MG_FP64 && precision != MIST_MG_MIXED) return fail(ctx, MIST_ERR_PARAM, "precision");
    if (!(tol > 0.0)) return fail(ctx, MIST_ERR_PARAM, "tol must be > 0");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    double mx[3];
    RET_IF(mist::launch_maxabs_uvw(ctx, mx));  // synchronises: the previous step is complete
    RET_IF(check_overflow(ctx, true));
    double dt = dt_from_max(ctx, mx);
    if (clamp_to > 0.0 && *t + dt > clamp_to) dt = clamp_to - *t;
    *t += dt;
    if (dt_used) *dt_used = dt;
    RET_IF(transport_ns_body(ctx, dt));
    RET_IF(mist_compute_band(ctx));
    RET_IF(mist::launch_props(ctx));
    RET_IF(mist::launch_interp_face(ctx, MIST_FIELD_US, MIST_FIELD_VS, MIST_FIELD_WS));
    const bool sl = ctx->p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN;
    if (!sl) RET_IF(mist::launch_pressure_rhs(ctx));  // the SL equation has no right-hand side
    int64_t cyc = 0;
    double res = 0.0;
    int st;
    {
        StageTimer tm(ctx, MIST_STAGE_MG);
        st = sl ? mist::pressure_solve_sl(ctx, dt, tol, &cyc, &res, nullptr)
                : mist::pressure_solve(ctx, dt, tol, (int)precision, &cyc, &res);
    }
    if (cycles) *cycles = cyc;
    if (final_resid) *final_resid = res;
    if (st && st != MIST_ERR_NOCONV) { g_last_error = ctx->err; return st; }
    RET_IF(mist::launch_corrector(ctx, dt));
    RET_IF(mist::launch_interp_center(ctx));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_U));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_V));
    RET_IF(mist::launch_borders(ctx, MIST_FIELD_W));
    RET_IF(mist::launch_divergence(ctx, dt));
    RET_IF(check_overflow(ctx));
    if (st) g_last_error = ctx->err;
    return st;
}

int mist_pressure_rhs(mist_ctx* ctx) {
    CHECK_CTX();
    RET_IF(mist::launch_pressure_rhs(ctx));
    return MIST_OK;
}

int mist_pressure_solve(mist_ctx* ctx, double dt, double tol, int64_t precision, int64_t* cycles, double* final_resid) {
    CHECK_CTX();
    if (precision != MIST_MG_FP64 && precision != MIST_MG_MIXED) return fail(ctx, MIST_ERR_PARAM, "precision");
    if (!(tol > 0.0) || !(dt > 0.0)) return fail(ctx, MIST_ERR_PARAM, "dt and tol must be > 0");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    int64_t cyc = 0;
    double res = 0.0;
    int st;
    {
        StageTimer tm(ctx, MIST_STAGE_MG);
        st = ctx->p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN ? mist::pressure_solve_sl(ctx, dt, tol, &cyc, &res, nullptr)
                                                               : mist::pressure_solve(ctx, dt, tol, (int)precision, &cyc, &res);
    }
    if (cycles) *cycles = cyc;
    if (final_resid) *final_resid = res;
    if (st) g_last_error = ctx->err;
    return st;
}

int mist_pressure_operator(mist_ctx* ctx, double dt) {
    CHECK_CTX();
    if (ctx->p.pressure_scheme != MIST_PS_SEMI_LAGRANGIAN)
        return fail(ctx, MIST_ERR_PARAM, "mist_pressure_operator: the semi-Lagrangian scheme's operator");
    if (ctx->p.xper || ctx->p.yper || ctx->p.zper) return fail(ctx, MIST_ERR_PARAM, "semi-Lagrangian pressure: periodic axes are not supported");
    if (!(dt > 0.0)) return fail(ctx, MIST_ERR_PARAM, "dt must be > 0");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::pressure_operator_sl(ctx, dt));
    return MIST_OK;
}

int mist_pressure_jacobian(mist_ctx* ctx, double dt, double* out, size_t n) {
    CHECK_CTX();
    if (ctx->p.pressure_scheme != MIST_PS_SEMI_LAGRANGIAN)
        return fail(ctx, MIST_ERR_PARAM, "mist_pressure_jacobian: the semi-Lagrangian scheme's operator");
    if (ctx->p.xper || ctx->p.yper || ctx->p.zper) return fail(ctx, MIST_ERR_PARAM, "semi-Lagrangian pressure: periodic axes are not supported");
    if (!(dt > 0.0) || !out || n != (size_t)(ctx->fn[MIST_FIELD_P][0] * ctx->fn[MIST_FIELD_P][1] * ctx->fn[MIST_FIELD_P][2]))
        return fail(ctx, MIST_ERR_PARAM, "mist_pressure_jacobian: dt > 0 and a cell-extent output array");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    RET_IF(mist::pressure_jacobian_sl(ctx, dt, out));
    return MIST_OK;
}

int mist_corrector(mist_ctx* ctx, double dt) {
    CHECK_CTX();
    RET_IF(mist::launch_corrector(ctx, dt));
    return MIST_OK;
}

int mist_reduce(mist_ctx* ctx, int kind, int field, double* out) {
    CHECK_CTX();
    if (!out || kind < 0 || kind > 3 || field < 0 || field >= MIST_FIELD_COUNT || field == MIST_FIELD_BAND ||
        field == MIST_FIELD_MASK || face_field(field))
        return fail(ctx, MIST_ERR_PARAM, "bad reduction");
    RET_IF(mist::launch_reduce(ctx, kind, field, out));
    if (ctx->defer_checks) RET_IF(check_overflow(ctx, true));
    return MIST_OK;
}

static int batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* s, double* out, int64_t m,
                 bool dist) {
    CHECK_CTX();
    if (m < 0 || (m > 0 && (!ijk || !nrm || !s || !out))) return fail(ctx, MIST_ERR_PARAM, "bad batch");
    if (m == 0) return MIST_OK;
    const mist::DevMesh& mm = ctx->m;
    for (int64_t q = 0; q < m; ++q) {
        const int64_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
        if (i < 1 - mm.ng || j < 1 - mm.ng || k < 1 - mm.ng || i > mm.Nx + mm.ng - (dist ? 0 : 1) ||
            j > mm.Ny + mm.ng - (dist ? 0 : 1) || k > mm.Nz + mm.ng - (dist ? 0 : 1))
            return fail(ctx, MIST_ERR_PARAM, "cell index out of range");
    }
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    int64_t* d_ijk = nullptr;
    double *d_n = nullptr, *d_s = nullptr, *d_o = nullptr;
    auto cleanup = [&]() {
        if (d_ijk) (void)hipFree(d_ijk);
        if (d_n) (void)hipFree(d_n);
        if (d_s) (void)hipFree(d_s);
        if (d_o) (void)hipFree(d_o);
    };
    if (hipMalloc(&d_ijk, m * 3 * sizeof(int64_t)) != hipSuccess || hipMalloc(&d_n, m * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_s, m * sizeof(double)) != hipSuccess || hipMalloc(&d_o, m * sizeof(double)) != hipSuccess) {
        cleanup();
        return fail(ctx, MIST_ERR_OOM, "batch allocation");
    }
    hipError_t e = hipMemcpyAsync(d_ijk, ijk, m * 3 * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_n, nrm, m * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_s, s, m * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    int st = MIST_OK;
    if (e == hipSuccess)
        st = dist ? mist::launch_plic_dist_batch(ctx, d_ijk, d_n, d_s, d_o, m)
                  : mist::launch_plic_vf_batch(ctx, d_ijk, d_n, d_s, d_o, m);
    if (e == hipSuccess && st == MIST_OK) e = hipMemcpyAsync(out, d_o, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return fail(ctx, MIST_ERR_HIP, hipGetErrorString(e));
    return st;
}

int mist_plic_dist_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* vf, double* dist_out,
                         int64_t m) {
    return batch(ctx, ijk, nrm, vf, dist_out, m, true);
}

int mist_plic_vf_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* dist, double* vf_out,
                       int64_t m) {
    return batch(ctx, ijk, nrm, dist, vf_out, m, false);
}

int mist_plic_dist_batch_noalloc(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* vf,
                                 double* dist_out, int64_t m) {
    CHECK_CTX();
    if (m < 0 || (m > 0 && (!ijk || !nrm || !vf || !dist_out))) return fail(ctx, MIST_ERR_PARAM, "bad batch");
    if (m > NOALLOC_BATCH_MAX)
        return fail(ctx, MIST_ERR_PARAM, "mist_plic_dist_batch_noalloc: batch of " + std::to_string(m) +
                                             " entries exceeds the limit of " + std::to_string(NOALLOC_BATCH_MAX));
    if (m == 0) return MIST_OK;
    const mist::DevMesh& mm = ctx->m;
    for (int64_t q = 0; q < m; ++q) {
        const int64_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
        if (i < 1 - mm.ng || j < 1 - mm.ng || k < 1 - mm.ng || i > mm.Nx + mm.ng || j > mm.Ny + mm.ng || k > mm.Nz + mm.ng)
            return fail(ctx, MIST_ERR_PARAM, "cell index out of range");
    }
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    double* base = (double*)g_scratch_ptr[ctx->device];  // set by mist_ctx_create on this device
    if (!base) return fail(ctx, MIST_ERR_STATE, "scratch address of the device is unknown");
    // staging inside the scratch array: [ijk 3m int64][nrm 3m][vf m][dist m]
    int64_t* d_ijk = (int64_t*)base;
    double* d_n = base + 3 * m;
    double* d_s = base + 6 * m;
    double* d_o = base + 7 * m;
    MIST_HIP_TRY(hipMemcpyAsync(d_ijk, ijk, m * 3 * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(d_n, nrm, m * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(d_s, vf, m * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RET_IF(mist::launch_plic_dist_batch(ctx, d_ijk, d_n, d_s, d_o, m));
    MIST_HIP_TRY(hipMemcpyAsync(dist_out, d_o, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MIST_OK;
}

// ---- adjoint ---------------------------------------------------------------------------
int mist_transport_vjp(mist_ctx* ctx, double dt) {
    CHECK_CTX();
    if (ctx->p.solveNS) return fail(ctx, MIST_ERR_PARAM, "mist_transport_vjp: the adjoint covers the VOF transport (solveNS = 0)");
    if (!(dt > 0.0)) return fail(ctx, MIST_ERR_PARAM, "dt must be > 0");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    const mist::DevMesh& m = ctx->m;
    if (!ctx->adj_lam) {
        const size_t nvof = (size_t)m.vnx * m.vny * m.vnz;
        const size_t ncell_ext = (size_t)m.cnx * m.cny * m.cnz;
        const size_t nv = (size_t)(m.Nx + 1) * (m.Ny + 1) * (m.Nz + 1);
        const size_t ncell = (size_t)m.Nx * m.Ny * m.Nz;
        if (hipMalloc(&ctx->adj_lam, nvof * sizeof(double)) != hipSuccess ||
            hipMalloc(&ctx->adj_plane, 4 * ncell_ext * sizeof(double)) != hipSuccess ||
            hipMalloc(&ctx->adj_pv, 3 * nv * sizeof(double)) != hipSuccess ||
            hipMalloc(&ctx->adj_sums, 6 * ncell * sizeof(double)) != hipSuccess)
            return fail(ctx, MIST_ERR_OOM, "adjoint work allocation");
    }
    // the forward state of the step: normals (and the ELVIRA choice kept with them), PLIC
    // constants and band of this VF, recomputed when stale
    RET_IF(mist_compute_normal(ctx));
    if (ctx->plic_version != ctx->vf_version) RET_IF(mist_compute_plic(ctx));
    if (!ctx->band_list_valid) RET_IF(mist_compute_band(ctx));
    if (ctx->prof.on) {
        RET_IF(mist::launch_transport_vjp(ctx, dt));  // profiled launches are timed one by one
    } else {
        // the ~25 launches of the reverse step replayed as one graph, captured for this dt: the
        // launches read the context's arrays through fixed pointers
        RET_IF(mist::vjp_prepare(ctx));
        if (ctx->vjp_exec && ctx->vjp_dt != dt) {
            (void)hipGraphExecDestroy(ctx->vjp_exec);
            (void)hipGraphDestroy(ctx->vjp_graph);
            ctx->vjp_exec = nullptr;
            ctx->vjp_graph = nullptr;
        }
        if (!ctx->vjp_exec) {
            MIST_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const int st = mist::launch_transport_vjp(ctx, dt);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            if (st) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            MIST_HIP_TRY(ec);
            ctx->vjp_graph = g;
            MIST_HIP_TRY(hipGraphInstantiate(&ctx->vjp_exec, g, nullptr, nullptr, 0));
            ctx->vjp_dt = dt;
        }
        MIST_HIP_TRY(hipGraphLaunch(ctx->vjp_exec, ctx->stream));
    }
    RET_IF(check_overflow(ctx));
    return MIST_OK;
}

// A checkpoint holds the state a step starts from: VF with its ghosts, the band flags and the
// band-cell list with its length, the normals, the PLIC constants and ELVIRA's choice per
// interface cell, each part at a 256-byte aligned offset of one allocation per slot.
struct CkptPart {
    void* dev;
    size_t bytes;
};
static int ckpt_parts(mist_ctx* ctx, CkptPart (&parts)[9], size_t& total) {
    const size_t ncell = (size_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    parts[0] = {ctx->f[MIST_FIELD_VOF], field_bytes(ctx, MIST_FIELD_VOF)};
    parts[1] = {ctx->f[MIST_FIELD_BAND], field_bytes(ctx, MIST_FIELD_BAND)};
    parts[2] = {ctx->f[MIST_FIELD_NX], field_bytes(ctx, MIST_FIELD_NX)};
    parts[3] = {ctx->f[MIST_FIELD_NY], field_bytes(ctx, MIST_FIELD_NY)};
    parts[4] = {ctx->f[MIST_FIELD_NZ], field_bytes(ctx, MIST_FIELD_NZ)};
    parts[5] = {ctx->f[MIST_FIELD_PLIC_D], field_bytes(ctx, MIST_FIELD_PLIC_D)};
    parts[6] = {ctx->elv_best, ncell * sizeof(int)};
    parts[7] = {ctx->band_list, ncell * sizeof(int)};
    parts[8] = {ctx->counters, sizeof(int)};  // the band-cell count
    total = 0;
    for (const CkptPart& q : parts) total += (q.bytes + 255) & ~(size_t)255;
    return MIST_OK;
}

// The parts of a checkpoint copied by one launch: 16-byte words, then the tail bytes (every
// part starts 256-byte aligned).
struct CopyTab {
    const char* src[9];
    char* dst[9];
    unsigned long long bytes[9];
};
__global__ void k_copy_parts(CopyTab t) {
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
    for (int p = 0; p < 9; ++p) {
        const size_t n16 = t.bytes[p] / 16;
        const uint4* s = (const uint4*)t.src[p];
        uint4* d = (uint4*)t.dst[p];
        for (size_t q = tid; q < n16; q += st) d[q] = s[q];
        for (size_t b = 16 * n16 + tid; b < t.bytes[p]; b += st) t.dst[p][b] = t.src[p][b];
    }
}

static int ckpt_copy(mist_ctx* ctx, int slot, bool save) {
    CHECK_CTX();
    if (slot < 0 || slot >= 4096) return fail(ctx, MIST_ERR_PARAM, "checkpoint slot must be 0..4095");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    CkptPart parts[9];
    size_t total = 0;
    RET_IF(ckpt_parts(ctx, parts, total));
    if ((size_t)slot >= ctx->ckpt.size()) ctx->ckpt.resize((size_t)slot + 1, nullptr);
    if (!ctx->ckpt[slot]) {
        if (!save) return fail(ctx, MIST_ERR_STATE, "checkpoint slot " + std::to_string(slot) + " is empty");
        if (hipMalloc(&ctx->ckpt[slot], total) != hipSuccess) return fail(ctx, MIST_ERR_OOM, "checkpoint allocation");
    }
    if (save) {
        // the band, normals and PLIC constants of this VF, where they are stale
        RET_IF(mist_compute_normal(ctx));
        if (ctx->plic_version != ctx->vf_version) RET_IF(mist_compute_plic(ctx));
        if (!ctx->band_list_valid) RET_IF(mist_compute_band(ctx));
    }
    char* base = (char*)ctx->ckpt[slot];
    size_t off = 0;
    CopyTab tab;
    for (int p = 0; p < 9; ++p) {
        char* slotp = base + off;
        tab.src[p] = save ? (const char*)parts[p].dev : slotp;
        tab.dst[p] = save ? slotp : (char*)parts[p].dev;
        tab.bytes[p] = parts[p].bytes;
        off += (parts[p].bytes + 255) & ~(size_t)255;
    }
    const int grid = (int)std::min<size_t>(2048, std::max<size_t>(1, total / (16 * 256)));
    hipLaunchKernelGGL(k_copy_parts, dim3(grid), dim3(256), 0, ctx->stream, tab);
    MIST_HIP_TRY(hipGetLastError());
    if (!save) {
        ctx->vf_version++;
        ctx->normals_version = ctx->vf_version;
        ctx->plic_version = ctx->vf_version;
        ctx->band_list_valid = true;
    }
    return MIST_OK;
}

int mist_checkpoint_save(mist_ctx* ctx, int slot) { return ckpt_copy(ctx, slot, true); }
int mist_checkpoint_restore(mist_ctx* ctx, int slot) { return ckpt_copy(ctx, slot, false); }

int mist_profile_enable(mist_ctx* ctx, int on) {
    CHECK_CTX();
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    KernelProfile& P = ctx->prof;
    if (on) {
        if (!P.d_units) MIST_HIP_TRY(hipMalloc(&P.d_units, MIST_PROF_COUNT * sizeof(unsigned long long)));
        hipLaunchKernelGGL(mist::k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, P.d_units, (int)MIST_PROF_COUNT);
        MIST_HIP_TRY(hipGetLastError());
        MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < MIST_PROF_COUNT; ++k) {
            for (auto& pr : P.pending[k]) { P.pool.push_back(pr.first); P.pool.push_back(pr.second); }
            P.pending[k].clear();
            P.ms[k] = 0.0;
            P.launches[k] = 0;
            P.host_units[k] = 0;
        }
    }
    P.on = on != 0;
    return MIST_OK;
}

int mist_profile_read(mist_ctx* ctx, int kernel, double* total_ms, int64_t* launches, int64_t* units) {
    CHECK_CTX();
    if (kernel < 0 || kernel >= MIST_PROF_COUNT) return fail(ctx, MIST_ERR_PARAM, "bad kernel id");
    MIST_HIP_TRY(hipSetDevice(ctx->device));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    KernelProfile& P = ctx->prof;
    for (int k = 0; k < MIST_PROF_COUNT; ++k) {
        for (auto& pr : P.pending[k]) {
            float ms = 0.f;
            MIST_HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
            P.ms[k] += ms;
            P.pool.push_back(pr.first);
            P.pool.push_back(pr.second);
        }
        P.pending[k].clear();
    }
    unsigned long long du[MIST_PROF_COUNT] = {};
    if (P.d_units) MIST_HIP_TRY(hipMemcpy(du, P.d_units, sizeof(du), hipMemcpyDeviceToHost));
    if (total_ms) *total_ms = P.ms[kernel];
    if (launches) *launches = P.launches[kernel];
    if (units) *units = (int64_t)du[kernel] + P.host_units[kernel];
    return MIST_OK;
}

int mist_stage_times(const mist_ctx* ctx, double* ms, int n) {
    if (!ctx || !ms) return MIST_ERR_PARAM;
    for (int q = 0; q < n && q < MIST_STAGE_COUNT; ++q) ms[q] = ctx->stage_ms[q];
    return MIST_OK;
}

}  // extern "C"
