// mgf_capi.cpp — the extern "C" boundary (include/mgf.h): contexts, scenes,
// the stages of CGridFluidSolver::update() and the operator seams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "mgf_internal.h"

using namespace mgf;

namespace mgf {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
static thread_local ProbeSink* g_probe_sink = nullptr;
static thread_local double g_probe_bytes = 0.0;
ProbeSink*& probe_sink() { return g_probe_sink; }
void probe_next_bytes(double bytes) { g_probe_bytes = bytes; }
double probe_take_bytes() {
    const double b = g_probe_bytes;
    g_probe_bytes = 0.0;
    return b;
}

bool ProbeSink::next_pair(hipEvent_t& a, hipEvent_t& b) {
    if (ev.size() < 2 * (used + 1)) {
        hipEvent_t x = nullptr, y = nullptr;
        if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) {
            (void)hipGetLastError();
            if (x) (void)hipEventDestroy(x);
            return false;
        }
        ev.push_back(x);
        ev.push_back(y);
    }
    a = ev[2 * used];
    b = ev[2 * used + 1];
    return true;
}

// Template instantiations are reported under the kernel's base name.
void ProbeSink::collect() {
    for (size_t k = 0; k < used; ++k) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        std::string nm(name[k]);
        const size_t lt = nm.find('<');
        if (lt != std::string::npos) nm.resize(lt);
        done_name.push_back(nm);
        done_ms.push_back(double(t));
        done_bytes.push_back(bytes[k]);
    }
    used = 0;
    name.clear();
    bytes.clear();
}

void ProbeSink::release() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    name.clear();
    bytes.clear();
    used = 0;
}
}  // namespace mgf

// Scratch of mgf_op_extrapolate_fc_noalloc (layout and limits with the entry
// point below): 35.1 MB held by every device the library runs on until the
// process ends.  Host code never passes this name to a kernel: it uses the
// device address looked up per device (na_setup).
static constexpr size_t kNoAllocScratchBytes = 35143424;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kPlainDotBlocks = 1024;
constexpr int kResidualBlocks = 1024;
constexpr int kDivBlocks = 1024;          // divergence check (grid-stride)

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

mgf_status fail(mgf_status s, const std::string& msg) {
    set_error(msg);
    return s;
}

// Device allocation owned by the scene; a dry scene only counts the bytes.
template <class T>
mgf_status dalloc(mgf_scene sc, T** out, size_t count) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (sc->dry) {
        sc->device_bytes += int64_t(bytes);
        *out = nullptr;
        return MGF_OK;
    }
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MGF_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    sc->allocations.push_back(p);
    sc->device_bytes += int64_t(bytes);
    *out = static_cast<T*>(p);
    return MGF_OK;
}

#define TRY(expr)                       \
    do {                                \
        mgf_status s_ = (expr);         \
        if (s_ != MGF_OK) return s_;    \
    } while (0)

Vel vel_of(const float* u, const float* v, const float* w) { return Vel{u, v, w}; }

bool finite3(const float* a) {
    return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]);
}

mgf_status validate_desc(const mgf_scene_desc* d) {
    if (!d) return fail(MGF_ERR_INVALID, "null descriptor");
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(MGF_ERR_INVALID, "resolution must be >= 1");
    const double cells = double(d->nx) * d->ny * d->nz;
    const double faces = double(d->nx + 1) * d->ny * d->nz + double(d->nx) * (d->ny + 1) * d->nz +
                         double(d->nx) * d->ny * (d->nz + 1);
    if (cells >= 2147483648.0 || faces >= 1073741824.0)
        return fail(MGF_ERR_INVALID, "grid too large (faces must be < 2^30)");
    for (int k = 0; k < 3; ++k)
        if (!(d->spacing[k] != 0.0f) || !std::isfinite(d->spacing[k]) || !std::isfinite(d->origin[k]))
            return fail(MGF_ERR_INVALID, "spacing must be finite and non-zero, origin finite");
    if (!std::isfinite(d->dt)) return fail(MGF_ERR_INVALID, "dt must be finite");
    if (!(d->density > 0.0f)) return fail(MGF_ERR_INVALID, "density must be > 0");
    if (!finite3(d->gravity) || !finite3(d->external_force))
        return fail(MGF_ERR_INVALID, "forces must be finite");
    if (d->rk_order < 1 || d->rk_order > 3) return fail(MGF_ERR_INVALID, "rk_order must be 1..3");
    if (d->tol_mode != MGF_TOL_ABS_SQRT_NORM && d->tol_mode != MGF_TOL_RELATIVE)
        return fail(MGF_ERR_INVALID, "unknown tol_mode");
    if (!(d->tolerance > 0.0)) return fail(MGF_ERR_INVALID, "tolerance must be > 0");
    if (d->precision != MGF_PRECISION_PLAIN && d->precision != MGF_PRECISION_MIXED)
        return fail(MGF_ERR_INVALID, "unknown precision");
    if (d->preconditioner != MGF_PRECOND_IDENTITY && d->preconditioner != MGF_PRECOND_JACOBI &&
        d->preconditioner != MGF_PRECOND_MULTIGRID)
        return fail(MGF_ERR_INVALID, "unknown preconditioner");
    if (d->precision == MGF_PRECISION_PLAIN && d->preconditioner != MGF_PRECOND_IDENTITY)
        return fail(MGF_ERR_UNSUPPORTED, "PLAIN precision runs the reference's identity-preconditioned CG only");
    if (d->mg_sweeps > 8) return fail(MGF_ERR_INVALID, "mg_sweeps must be <= 8");
    if (!(d->mg_omega < 2.0f) || !std::isfinite(d->mg_omega)) return fail(MGF_ERR_INVALID, "mg_omega must be < 2");
    if (d->mg_kernels != MGF_MG_AUTO && d->mg_kernels != MGF_MG_PER_SWEEP && d->mg_kernels != MGF_MG_FUSED)
        return fail(MGF_ERR_INVALID, "unknown mg_kernels");
    if (d->extrapolation_layers < 0 || d->extrapolation_layers > 1000000)
        return fail(MGF_ERR_INVALID, "extrapolation_layers out of range");
    return MGF_OK;
}

size_t field_count(const mgf_scene_s* s, int field) {
    switch (field) {
        case MGF_FIELD_VEL_X: case MGF_FIELD_SOLID_VEL_X: return s->nfx;
        case MGF_FIELD_VEL_Y: case MGF_FIELD_SOLID_VEL_Y: return s->nfy;
        case MGF_FIELD_VEL_Z: case MGF_FIELD_SOLID_VEL_Z: return s->nfz;
        case MGF_FIELD_MATRIX_A: return 4 * s->ncell;
        default: return s->ncell;
    }
}

void scales(const mgf_scene_s* s, float sa[3], float sg[3]) {
    const float dt = s->desc.dt, rho = s->desc.density;
    const float h[3] = {s->desc.spacing[0], s->desc.spacing[1], s->desc.spacing[2]};
    for (int k = 0; k < 3; ++k) {
        // CPressureSolver::__buildMatrix / __applyPressureGradient
        // (MPhysics/PressureSolver.cpp:148-164), evaluated in float.
        sa[k] = dt / (rho * h[k] * h[k]);
        sg[k] = dt / (rho * h[k]);
    }
}

// ---- kernel probes: the scene's sink is installed for the host thread while
// one of its entry points runs (ProbeSink, mgf_internal.h).
struct ProbeScope {
    ProbeSink* prev;
    explicit ProbeScope(mgf_scene s) : prev(probe_sink()) { probe_sink() = s->probes ? &s->sink : nullptr; }
    ~ProbeScope() { probe_sink() = prev; }
};

mgf_status sync(mgf_scene s) {
    MGF_HIP(hipStreamSynchronize(s->stream));
    if (s->sink.used) s->sink.collect();
    return MGF_OK;
}

// Host-side sum of FP64 partials (fixed order).
double host_sum(const double* p, int n, int stride, int k) {
    double acc = 0;
    for (int i = 0; i < n; ++i) acc += p[size_t(i) * stride + k];
    return acc;
}

mgf_status dot_host(mgf_scene s, const float* a, const float* b, size_t n, double* out) {
    hipStream_t st = s->stream;
    MGF_HIP(launch_dot_partials(st, a, b, n, s->red.partials, kPlainDotBlocks));
    MGF_HIP(hipMemcpyAsync(s->host_scratch, s->red.partials, sizeof(double) * kPlainDotBlocks,
                           hipMemcpyDeviceToHost, st));
    MGF_HIP(hipStreamSynchronize(st));
    *out = host_sum(s->host_scratch, kPlainDotBlocks, 1, 0);
    return MGF_OK;
}

// ---------------------------------------------------------------------------
// Boundaries: CBoundarys::updateBoundarys (MPhysics/Boundary.cpp:115-150).
mgf_status rebuild_boundaries(mgf_scene s);
mgf_status refresh_boundaries(mgf_scene s) {
    bool moving = false;
    for (auto& so : s->solids)
        if (so.vel[0] != 0.0f || so.vel[1] != 0.0f || so.vel[2] != 0.0f) moving = true;
    // __moveBoundarys: translation += -dt * velocity (cuBLAS axpy, no skip)
    for (auto& so : s->solids)
        for (int k = 0; k < 3; ++k) so.trans[k] = std::fma(-s->desc.dt, so.vel[k], so.trans[k]);
    if (!s->boundary_dirty && !moving) return MGF_OK;
    return rebuild_boundaries(s);
}

// The solid marker, solid mask and solid face velocities from the solids at
// their current translations.
mgf_status rebuild_boundaries(mgf_scene s) {
    hipStream_t st = s->stream;
    // zero-filled marker, then each solid in order
    MGF_HIP(launch_fill(st, s->solid_total, s->ncell, 0.0f));
    float* tmp = s->ap;
    std::vector<float> vel3;
    for (size_t k = 0; k < s->solids.size(); ++k) {
        MGF_HIP(launch_resample_solid(st, s->g, s->solids[k].sdf0, s->solids[k].trans, tmp));
        MGF_HIP(launch_solids_marker(st, s->ncell, tmp, int(k) + 1, s->solid_total));
        for (int q = 0; q < 3; ++q) vel3.push_back(s->solids[k].vel[q]);
    }
    MGF_HIP(launch_solid_mask(st, s->ncell, s->solid_total, s->solid_mask));
    if (vel3.empty()) vel3.assign(3, 0.0f);
    MGF_HIP(launch_solids_vel(st, s->g, s->solid_total, vel3.data(), int(s->solids.size()), s->us,
                              s->vs, s->ws));
    s->boundary_dirty = false;
    s->solids_unbuilt = false;
    return MGF_OK;
}

// ---------------------------------------------------------------------------
// _computeAdvection + (optionally) _computeExternalForces.
mgf_status advect(mgf_scene s, bool with_forces) {
    hipStream_t st = s->stream;
    const int c = s->cur, n = 1 - s->cur;
    const float dt = s->desc.dt;
    float g[3], f[3];
    int addg = 0, addf = 0;
    if (with_forces) {
        for (int k = 0; k < 3; ++k) {
            g[k] = s->desc.gravity[k] * dt;
            f[k] = s->desc.external_force[k] * dt;
        }
        auto on = [](const float* a) {
            return !(std::fabs(a[0]) < MGF_FLT_EPSILON && std::fabs(a[1]) < MGF_FLT_EPSILON &&
                     std::fabs(a[2]) < MGF_FLT_EPSILON);
        };
        addg = on(g);
        addf = on(f);
    }
    Vel V = vel_of(s->u[c], s->v[c], s->w[c]);
    probe_next_bytes(12.0 * double(s->nfx + s->nfy + s->nfz));
    MGF_HIP(launch_advect_fc(st, s->g, V, V, dt, s->desc.rk_order, s->u[n], s->v[n], s->w[n], addg,
                             g, addf, f));
    if (s->color[0][0]) {
        const float* ic[3] = {s->color[c][0], s->color[c][1], s->color[c][2]};
        float* oc[3] = {s->color[n][0], s->color[n][1], s->color[n][2]};
        probe_next_bytes(32.0 * double(s->ncell));
        MGF_HIP(launch_advect_cc4(st, s->g, s->phi[c], ic, V, dt, s->desc.rk_order, s->phi[n], oc));
    } else {
        probe_next_bytes(8.0 * double(s->ncell));
        MGF_HIP(launch_advect_cc(st, s->g, s->phi[c], V, dt, s->desc.rk_order, s->phi[n]));
    }
    s->cur = n;
    return MGF_OK;
}

mgf_status apply_forces(mgf_scene s) {
    hipStream_t st = s->stream;
    const float dt = s->desc.dt;
    const float* incs[2] = {s->desc.gravity, s->desc.external_force};
    for (int q = 0; q < 2; ++q) {
        float a[3] = {incs[q][0] * dt, incs[q][1] * dt, incs[q][2] * dt};
        if (std::fabs(a[0]) < MGF_FLT_EPSILON && std::fabs(a[1]) < MGF_FLT_EPSILON &&
            std::fabs(a[2]) < MGF_FLT_EPSILON)
            continue;
        MGF_HIP(launch_add_uniform(st, s->u[s->cur], s->nfx, a[0]));
        MGF_HIP(launch_add_uniform(st, s->v[s->cur], s->nfy, a[1]));
        MGF_HIP(launch_add_uniform(st, s->w[s->cur], s->nfz, a[2]));
    }
    return MGF_OK;
}

// ---------------------------------------------------------------------------
// Plain CG: CCuMatrixFreePCG::solvePCGInvDia (MPhysics/CuMatrixFreePCG.cpp:
// 22-65) with the identity preconditioner, scalars on the host.
mgf_status cg_plain(mgf_scene s, const float sa[3], int64_t max_it, mgf_step_stats* acc) {
    hipStream_t st = s->stream;
    const size_t n = s->ncell;
    const auto& d = s->desc;
    double bb;
    TRY(dot_host(s, s->b, s->b, n, &bb));
    const float bn = float(std::sqrt(bb));
    auto stop = [&](float rn) {
        if (d.tol_mode == MGF_TOL_ABS_SQRT_NORM) return std::sqrt(rn) < float(d.tolerance);
        return double(rn) <= d.tolerance * double(bn);
    };
    MGF_HIP(launch_fill(st, s->x, n, 0.0f));
    s->x_is_64 = false;
    int64_t it = 0;
    bool conv = false;
    double rn_last = bn;
    if (d.tol_mode == MGF_TOL_ABS_SQRT_NORM ? std::sqrt(bn) < float(d.tolerance) : bn == 0.0f) {
        conv = true;
    } else {
        // r = A*0 = 0 -> scale(-1) -> += b, i.e. r = b; z = r; p = z.
        MGF_HIP(hipMemcpyAsync(s->r, s->b, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        MGF_HIP(hipMemcpyAsync(s->p0, s->r, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        // r . z with z = r; the value carries over from the previous
        // iteration's r . z (the same vector, the same sum).
        double t;
        TRY(dot_host(s, s->r, s->r, n, &t));
        float rz = float(t);
        for (; it < max_it;) {
            MGF_HIP(launch_matvec(st, s->g, sa, s->markers, s->p0, s->ap));
            ++it;
            TRY(dot_host(s, s->p0, s->ap, n, &t));
            const float pap = float(t);
            const float alpha = rz / pap;
            if (!(std::fabs(alpha) < MGF_FLT_EPSILON)) {
                MGF_HIP(launch_axpy(st, alpha, s->p0, s->x, n));
                MGF_HIP(launch_axpy(st, -alpha, s->ap, s->r, n));
            }
            TRY(dot_host(s, s->r, s->r, n, &t));
            const float rn = float(std::sqrt(t));
            rn_last = rn;
            if (stop(rn)) { conv = true; break; }
            if (!std::isfinite(rn)) break;   // reported as non-finite below
            const float rz_new = float(t);
            const float beta = rz_new / rz;
            rz = rz_new;
            MGF_HIP(launch_scal_add(st, beta, s->p0, s->r, n));
        }
    }
    // FP64 true residual of the FP32 pressure
    MGF_HIP(launch_true_residual32(st, s->g, sa, s->markers, s->x, s->b, s->red.partials,
                                   kResidualBlocks));
    MGF_HIP(hipMemcpyAsync(s->host_scratch, s->red.partials, sizeof(double) * 2 * kResidualBlocks,
                           hipMemcpyDeviceToHost, st));
    MGF_HIP(hipStreamSynchronize(st));
    const double rr = host_sum(s->host_scratch, kResidualBlocks, 2, 0);
    const double b2 = host_sum(s->host_scratch, kResidualBlocks, 2, 1);
    acc->cg_iterations += it;
    acc->converged = acc->converged && conv;
    // the same stop rule on the FP64 residual of the FP32 pressure
    const double rt = std::sqrt(rr), bt = std::sqrt(b2);
    const bool true_ok = d.tol_mode == MGF_TOL_ABS_SQRT_NORM ? std::sqrt(rt) < d.tolerance
                                                             : (bt == 0.0 ? rt == 0.0 : rt <= d.tolerance * bt);
    acc->true_converged = acc->true_converged && true_ok;
    acc->residual_rel = b2 > 0 ? std::sqrt(rr / b2) : std::sqrt(rr);
    acc->residual_recurrence = rn_last;
    acc->b_norm = std::sqrt(b2);
    acc->refinement_sweeps = 0;
    return MGF_OK;
}

// ---------------------------------------------------------------------------
// Preconditioners of the MIXED solver.
constexpr int kCoarsestSweeps = 8;
// the hierarchy stops once no axis exceeds this many cells (measured at 256^3
// and 512^3: 4, 8 give the same CG iterations, 8 one level fewer; 16 and 32
// add iterations)
constexpr int kMgStopExtent = 8;
constexpr size_t kCoarseCells = 8192;    // levels this small run in one work-group
constexpr int kCoarseLevelsMax = 8;
constexpr size_t kCoarseLdsMax = 131072;
// MGF_MG_AUTO fuses levels at least this large (measured at 256^3: fusing the
// 129^3 level too is 0.8% faster per step, the 65^3 one slower)
constexpr size_t kMarchMinCells = size_t(1) << 20;

// Whether level l (dims f; the next coarser level c) runs the fused descent
// and ascent: three sweeps, a 3-D level above the coarse cycle, under
// MGF_MG_AUTO at least kMarchMinCells (on smaller levels the march's serial
// steps cost more than the launches they save); on level 0 the ascent's
// partials of r . z must fit the scene's buffer.
// Blocks of a fused march resident at once: two per CU.
uint32_t march_slots(mgf_scene s) { return 2u * uint32_t(std::max(1, s->ctx->cu_count)); }

bool march_on(mgf_scene s, int l, const MgDims& f, const MgDims& c) {
    const auto& d = s->desc;
    const int nu = d.mg_sweeps > 0 ? d.mg_sweeps : 3;
    const int L = 1 + int(s->mg.size());
    const int l0 = s->mg_l0 < L ? s->mg_l0 : L;
    if (d.preconditioner != MGF_PRECOND_MULTIGRID || nu != 3 || d.mg_kernels == MGF_MG_PER_SWEEP) return false;
    if (l >= l0 || l + 1 >= L || !mg_march_ok(f)) return false;
    const size_t n = size_t(f.nx) * f.ny * f.nz;
    if (d.mg_kernels == MGF_MG_AUTO && n < kMarchMinCells) return false;
    const size_t part_cap = size_t(std::max<uint32_t>(kMaxBlocks, col_blocks(s->g))) * 2;
    return l > 0 || mg_march_parts(f, c) <= part_cap;
}

mgf_status ensure_sched(mgf_scene s);

mgf_status ensure_precond(mgf_scene s) {
    if (!s->z) {
        TRY(dalloc(s, &s->z, s->ncell));
        TRY(dalloc(s, &s->zt, s->ncell));
    }
    if (s->desc.preconditioner != MGF_PRECOND_MULTIGRID || s->mg_nx >= 0) return ensure_sched(s);
    // hierarchy: halve every axis longer than one cell until none exceeds
    // kMgStopExtent
    int nx = s->g.nx, ny = s->g.ny, nz = s->g.nz;
    while (std::max(nx, std::max(ny, nz)) > kMgStopExtent && s->mg.size() < 16) {
        MgLevelBuf L;
        L.d.nx = nx > 1 ? (nx + 1) / 2 : 1;
        L.d.ny = ny > 1 ? (ny + 1) / 2 : 1;
        L.d.nz = nz > 1 ? (nz + 1) / 2 : 1;
        const size_t n = size_t(L.d.nx) * L.d.ny * L.d.nz;
        TRY(dalloc(s, &L.markers, n));
        TRY(dalloc(s, &L.codes, n));
        TRY(dalloc(s, &L.b, n));
        TRY(dalloc(s, &L.x, n));
        TRY(dalloc(s, &L.t, n));
        s->mg.push_back(L);
        nx = L.d.nx; ny = L.d.ny; nz = L.d.nz;
    }
    s->mg_nx = nx; s->mg_ny = ny; s->mg_nz = nz;
    // the first level (below the pressure grid) from which the rest of the
    // hierarchy fits one work-group's LDS
    const int L = 1 + int(s->mg.size());
    s->mg_l0 = L;
    const size_t lds_cap = std::min<size_t>(kCoarseLdsMax, s->ctx->lds_per_block);
    for (int l = 1; l < L; ++l) {
        const MgDims& d = s->mg[size_t(l - 1)].d;
        if (size_t(d.nx) * d.ny * d.nz > kCoarseCells || L - l > kCoarseLevelsMax) continue;
        std::vector<MgDims> dims;
        for (int k = l; k < L; ++k) dims.push_back(s->mg[size_t(k - 1)].d);
        if (mg_coarse_lds_bytes(dims.data(), int(dims.size())) <= lds_cap) {
            s->mg_l0 = l;
            break;
        }
    }
    return ensure_sched(s);
}

// The schedule of each level fused now (mg_kernels may change between
// solves), allocated once.
mgf_status ensure_sched(mgf_scene s) {
    MgDims f{s->g.nx, s->g.ny, s->g.nz, 1.0f, 1.0f, 1.0f};
    for (size_t l = 0; l < s->mg.size(); ++l) {
        const MgDims& c = s->mg[l].d;
        uint32_t** sched = l == 0 ? &s->sched0 : &s->mg[l - 1].sched;
        size_t* sched_n = l == 0 ? &s->sched0_n : &s->mg[l - 1].sched_n;
        if (!*sched && march_on(s, int(l), f, c)) {
            *sched_n = mg_march_sched_words(f, c);
            TRY(dalloc(s, sched, *sched_n));
        }
        f = c;
    }
    return MGF_OK;
}

// Coarse markers and stencil codes of every level for the current markers;
// the operator scale of an axis drops by 4 where the axis is halved.
mgf_status mg_setup(mgf_scene s, const float sa[3]) {
    hipStream_t st = s->stream;
    MgDims f{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    const uint8_t* mf = s->markers;
    for (auto& L : s->mg) {
        L.d.sx = f.nx > 1 ? f.sx * 0.25f : f.sx;
        L.d.sy = f.ny > 1 ? f.sy * 0.25f : f.sy;
        L.d.sz = f.nz > 1 ? f.sz * 0.25f : f.sz;
        // a coarse cell with any air child is air (Dirichlet), else fluid if
        // any child is fluid, else solid
        MGF_HIP(launch_mg_coarsen(st, f, L.d, mf, L.markers, 0));
        MGF_HIP(launch_codes(st, L.d.nx, L.d.ny, L.d.nz, L.markers, L.codes));
        f = L.d;
        mf = L.markers;
    }
    // the fused levels' schedules, for these codes
    f = MgDims{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    const uint16_t* cf = s->codes;
    for (size_t l = 0; l < s->mg.size(); ++l) {
        const MgDims& c = s->mg[l].d;
        uint32_t* sched = l == 0 ? s->sched0 : s->mg[l - 1].sched;
        const size_t sched_n = l == 0 ? s->sched0_n : s->mg[l - 1].sched_n;
        if (sched && sched_n >= mg_march_sched_words(f, c)) {
            MGF_HIP(launch_mg_march_wet(st, f, c, cf, sched));
            MGF_HIP(launch_mg_march_plan(st, f, c, sched, march_slots(s)));
        }
        f = c;
        cf = s->mg[l].codes;
    }
    return MGF_OK;
}

// z = M^-1 r; *zout receives the buffer that holds z.  The last level-0
// kernel also reduces r . z and sets beta (see k_mg_smooth).
mgf_status apply_precond(mgf_scene s, const float sa[3], const float** zout) {
    hipStream_t st = s->stream;
    const auto& d = s->desc;
    const MgDims d0{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    if (d.preconditioner == MGF_PRECOND_JACOBI) {
        MGF_HIP(launch_mg_smooth(st, d0, s->codes, s->r, s->r, s->z, 1.0f, s->scal, s->red, 1, 1));
        *zout = s->z;
        return MGF_OK;
    }
    const int nu = d.mg_sweeps > 0 ? d.mg_sweeps : 3;
    const float omega = d.mg_omega > 0.0f ? d.mg_omega : 0.9f;
    const int L = 1 + int(s->mg.size());
    struct Lv {
        MgDims d; const uint16_t* codes; const float* b; float* x; float* t; const uint32_t* sched; double cells;
    };
    Lv lv[17];
    lv[0] = Lv{d0, s->codes, s->r, s->z, s->zt, s->sched0, s->wet0_cells};
    for (int l = 1; l < L; ++l) {
        auto& m = s->mg[size_t(l - 1)];
        lv[l] = Lv{m.d, m.codes, m.b, m.x, m.t, m.sched, m.march_cells};
    }
    const int skip = s->skip_dry ? 1 : 0;
    // levels [l0, L) run in one work-group
    const int l0 = s->mg_l0 < L ? s->mg_l0 : L;
    // descent: pre-smoothing from zero (the first two sweeps in one pass),
    // restriction of the residual
    // fused descent / ascent on the 3-D levels (AUTO: the large ones, where
    // they beat one kernel per sweep; on small levels the march's serial
    // steps cost more than the launches they save); the level-0 ascent
    // writes one partial of r . z per block into the scene's partials buffer
    auto fused = [&](int l) { return l + 1 < L && march_on(s, l, lv[l].d, lv[l + 1].d); };
    for (int l = 0; l < l0; ++l) {
        Lv& v = lv[l];
        const bool coarsest = l == L - 1;
        if (fused(l)) {
            // blocks without fluid store x = 0 and the coarse b = 0 only
            const double nl = double(v.d.nx) * v.d.ny * v.d.nz;
            const double wet = (skip && v.cells >= 0.0) ? v.cells : nl;
            probe_next_bytes(10.5 * wet + 4.5 * (nl - wet));
            MGF_HIP(launch_mg_descend(st, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].codes,
                                      const_cast<float*>(lv[l + 1].b), omega, s->scal, v.sched, skip));
            continue;
        }
        const int sweeps = coarsest ? kCoarsestSweeps : nu;
        auto rz_at = [&](int k) { return (l == 0 && coarsest && k == sweeps - 1) ? 1 : 0; };
        int k;
        const double ncell = double(v.d.nx) * v.d.ny * v.d.nz;
        if (sweeps >= 2) {
            probe_next_bytes(10.0 * ncell);
            MGF_HIP(launch_mg_smooth2_first(st, v.d, v.codes, v.b, v.x, omega, s->scal, s->red, rz_at(1)));
            k = 2;
        } else {
            MGF_HIP(launch_mg_smooth(st, v.d, v.codes, v.b, v.b, v.x, omega, s->scal, s->red, 1, rz_at(0)));
            k = 1;
        }
        for (; k < sweeps; ++k) {
            probe_next_bytes(14.0 * ncell);
            MGF_HIP(launch_mg_smooth(st, v.d, v.codes, v.b, v.x, v.t, omega, s->scal, s->red, 0, rz_at(k)));
            std::swap(v.x, v.t);
        }
        if (!coarsest) {
            probe_next_bytes(10.5 * ncell);
            MGF_HIP(launch_mg_restrict(st, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].codes,
                                       const_cast<float*>(lv[l + 1].b), s->scal));
        }
    }
    if (l0 < L) {
        MgDims dims[kCoarseLevelsMax];
        const uint16_t* codes[kCoarseLevelsMax];
        for (int l = l0; l < L; ++l) {
            dims[l - l0] = lv[l].d;
            codes[l - l0] = lv[l].codes;
        }
        MGF_HIP(launch_mg_coarse_cycle(st, dims, codes, L - l0, lv[l0].b, lv[l0].x, nu, kCoarsestSweeps, omega,
                                       s->scal));
    }
    // ascent: the first post-sweep applies the prolongation (injection) on
    // the fly, then the remaining sweeps
    for (int l = std::min(L - 2, l0 - 1); l >= 0; --l) {
        Lv& v = lv[l];
        auto rz_at = [&](int k) { return (l == 0 && k == nu - 1) ? 1 : 0; };
        const double ncell = double(v.d.nx) * v.d.ny * v.d.nz;
        if (fused(l)) {
            const double wet = (skip && v.cells >= 0.0) ? v.cells : ncell;
            probe_next_bytes(14.5 * wet + 4.0 * (ncell - wet));
            MGF_HIP(launch_mg_ascend(st, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].x, v.t, omega, s->scal,
                                     s->red, l == 0 ? 1 : 0, v.sched, skip));
            std::swap(v.x, v.t);
            continue;
        }
        probe_next_bytes(14.5 * ncell);
        MGF_HIP(launch_mg_smooth_prolong(st, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].x, v.t, omega, s->scal,
                                         s->red, rz_at(0)));
        std::swap(v.x, v.t);
        for (int k = 1; k < nu; ++k) {
            probe_next_bytes(14.0 * ncell);
            MGF_HIP(launch_mg_smooth(st, v.d, v.codes, v.b, v.x, v.t, omega, s->scal, s->red, 0, rz_at(k)));
            std::swap(v.x, v.t);
        }
    }
    *zout = lv[0].x;
    return MGF_OK;
}

// Mixed CG: FP32 vectors in two fused column kernels per iteration (plus the
// preconditioner), FP64 dots and scalars on the device, FP64 pressure
// accumulated over refinement sweeps whose residual b - A x is formed in FP64.
mgf_status cg_mixed(mgf_scene s, const float sa[3], int64_t max_it, mgf_step_stats* acc) {
    hipStream_t st = s->stream;
    const auto& d = s->desc;
    const bool ident = d.preconditioner == MGF_PRECOND_IDENTITY;
    const double inner = d.inner_tolerance > 0 ? d.inner_tolerance : 1e-3;
    const int max_sweeps = d.max_refinements > 0 ? d.max_refinements : 8;
    // iterations launched between two reads of the scalars: predicted from the
    // observed contraction of r.r (converged iterations still cost their
    // launches), bounded by `chunk`
    const int chunk = d.preconditioner == MGF_PRECOND_MULTIGRID ? 16 : (s->ncell < (1u << 20) ? 32 : 16);
    const int first_chunk = d.preconditioner == MGF_PRECOND_MULTIGRID ? 4 : chunk;
    // The estimate is rounded to the nearest count: an iteration launched past
    // convergence costs its whole V-cycle of (early-returning) launches, while
    // one too few costs only another read of the scalars.
    double log_rho = 0.0;   // log of the per-iteration factor of r.r, 0 = unknown
    auto predict = [&](double rr_now, double target) {
        if (!(log_rho < 0.0) || !(rr_now > target) || !(target > 0.0)) return first_chunk;
        const double n = std::floor(std::log(target / rr_now) / log_rho + 0.5);
        return int(std::max(1.0, std::min(double(chunk), n)));
    };
    MGF_HIP(launch_codes(st, s->g.nx, s->g.ny, s->g.nz, s->markers, s->codes));
    // allocated with the scene when it was created MIXED; a scene switched to
    // MIXED by mgf_scene_set_params gets it here
    if (!s->wet_flat) TRY(dalloc(s, &s->wet_flat, col_blocks(s->g)));
    MGF_HIP(launch_flat_wet(st, s->g, s->codes, s->wet_flat));
    const uint8_t* wet_flat = s->skip_dry ? s->wet_flat : nullptr;
    if (!ident) {
        TRY(ensure_precond(s));
        TRY(mg_setup(s, sa));
    }
    s->wet_flat_cells = s->wet0_cells = -1.0;
    if (s->probes && s->skip_dry) {
        // the bytes probed launches move: blocks without fluid move (almost) nothing
        std::vector<uint8_t> h(col_blocks(s->g));
        MGF_HIP(hipMemcpyAsync(h.data(), s->wet_flat, h.size(), hipMemcpyDeviceToHost, st));
        MGF_HIP(hipStreamSynchronize(st));
        s->wet_flat_cells = flat_wet_cells(s->g, h.data());
        // the cells each fused level's schedule marches
        MgDims f{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
        for (size_t l = 0; l < s->mg.size(); ++l) {
            const uint32_t* sched = l == 0 ? s->sched0 : s->mg[l - 1].sched;
            double* cells = l == 0 ? &s->wet0_cells : &s->mg[l - 1].march_cells;
            const MgDims& c = s->mg[l].d;
            *cells = -1.0;
            if (sched) {
                std::vector<uint32_t> hs(l == 0 ? s->sched0_n : s->mg[l - 1].sched_n);
                MGF_HIP(hipMemcpyAsync(hs.data(), sched, hs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
                MGF_HIP(hipStreamSynchronize(st));
                *cells = mg_march_sched_cells(f, c, hs.data());
            }
            f = c;
        }
    }
    const double ncell_all = double(s->ncell);
    const double flat_cells = s->wet_flat_cells >= 0.0 ? s->wet_flat_cells : ncell_all;
    bool first_pap = true;
    // x64 = 0 (or, warm, the previous pressure on the fluid cells and its
    // residual), ||b||^2 and the overall target
    const bool warm = d.warm_start && s->x64_valid;
    s->x64_valid = false;
    MGF_HIP(launch_refine_true(st, s->g, sa, s->codes, nullptr, s->x64, s->b, warm ? s->r : nullptr, warm ? 3 : 0,
                               s->red, s->red.out2));
    MGF_HIP(hipMemcpyAsync(s->host_scratch, s->red.out2, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    MGF_HIP(hipStreamSynchronize(st));
    const double b2 = s->host_scratch[1];
    double target2;
    if (d.tol_mode == MGF_TOL_ABS_SQRT_NORM) {
        const double t = d.tolerance * d.tolerance;  // sqrt(||r||) < tol
        target2 = t * t;
    } else {
        target2 = d.tolerance * d.tolerance * b2;
    }
    s->x_is_64 = true;
    int64_t it_total = 0;
    int sweeps = 0;
    double rr_true = warm ? s->host_scratch[0] : b2;
    bool conv = rr_true <= target2 || b2 == 0.0;
    const float* rsrc = warm ? s->r : s->b;
    CgScalars hs{};
    while (!conv && sweeps < max_sweeps && it_total < max_it) {
        MGF_HIP(launch_cg_start(st, s->g, rsrc, s->r, s->x, s->scal, s->red, inner, target2));
        first_pap = true;
        const float* zv = s->r;
        if (!ident) TRY(apply_precond(s, sa, &zv));
        // r.r at the sweep start and the sweep's target, as the device sets them
        const double rr0 = rr_true;
        const double sweep_target = std::max(inner * inner * rr0, target2);
        double rr_now = rr0;
        int done_before = 0;
        // p and A p (p0, p1) are updated in place by the fused kernels
        for (;;) {
            const int64_t left = max_it - it_total - done_before;
            const int k = int(std::min<int64_t>(predict(rr_now, sweep_target), std::max<int64_t>(left, 1)));
            for (int q = 0; q < k; ++q) {
                probe_next_bytes(22.0 * (first_pap ? ncell_all : flat_cells));
                MGF_HIP(launch_col_p_ap(st, s->g, sa, s->codes, zv, s->p0, s->p1, s->scal, s->red, wet_flat));
                first_pap = false;
                probe_next_bytes(24.0 * flat_cells);
                MGF_HIP(launch_col_update(st, s->g, s->p0, s->p1, s->x, s->r, s->scal, s->red, ident ? 1 : 0,
                                          wet_flat));
                if (!ident) TRY(apply_precond(s, sa, &zv));
            }
            MGF_HIP(hipMemcpyAsync(&hs, s->scal, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
            MGF_HIP(hipStreamSynchronize(st));
            if (hs.iters > 0 && hs.rr > 0.0 && rr0 > 0.0 && hs.rr < rr0)
                log_rho = std::log(hs.rr / rr0) / double(hs.iters);
            rr_now = hs.rr;
            done_before = hs.iters;
            if (hs.done || it_total + hs.iters >= max_it) break;
        }
        it_total += hs.iters;
        if (hs.done == 2) { rr_true = std::nan(""); break; }
        // x64 (+)= x; r = b - A x64 in FP64
        MGF_HIP(launch_refine_true(st, s->g, sa, s->codes, s->x, s->x64, s->b, s->r, (sweeps > 0 || warm) ? 2 : 1,
                                   s->red, s->red.out2));
        MGF_HIP(hipMemcpyAsync(s->host_scratch, s->red.out2, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
        MGF_HIP(hipStreamSynchronize(st));
        rr_true = s->host_scratch[0];
        ++sweeps;
        rsrc = s->r;
        if (rr_true <= target2) conv = true;
        if (!std::isfinite(rr_true)) break;
    }
    s->x64_valid = conv && std::isfinite(rr_true);
    acc->cg_iterations += it_total;
    acc->refinement_sweeps += sweeps;
    acc->converged = acc->converged && conv;
    acc->true_converged = acc->true_converged && conv;
    acc->residual_rel = b2 > 0 ? std::sqrt(rr_true / b2) : std::sqrt(rr_true);
    acc->residual_recurrence = std::sqrt(hs.rr);
    acc->b_norm = std::sqrt(b2);
    return MGF_OK;
}

// With particles, the markers come from their positions
// (CPressureSolver::solvePressure, PressureSolver.cpp:109-134).
mgf_status pressure(mgf_scene s, mgf_step_stats* acc, mgf_particles parts = nullptr) {
    hipStream_t st = s->stream;
    const int c = s->cur;
    float sa[3], sg[3];
    scales(s, sa, sg);
    if (parts) MGF_HIP(launch_particle_markers_u8(st, s->g, s->solid_mask, parts->n, parts->pos, s->markers));
    else {
        probe_next_bytes(6.0 * double(s->ncell));
        MGF_HIP(launch_markers(st, s->g, s->solid_mask, s->phi[c], s->markers));
    }
    probe_next_bytes(41.0 * double(s->ncell));
    MGF_HIP(launch_rhs(st, s->g, vel_of(s->u[c], s->v[c], s->w[c]), vel_of(s->us, s->vs, s->ws),
                       s->markers, s->b));
    const int64_t max_it = s->desc.max_iterations > 0 ? s->desc.max_iterations : int64_t(s->ncell);
    if (s->desc.precision == MGF_PRECISION_PLAIN) TRY(cg_plain(s, sa, max_it, acc));
    else TRY(cg_mixed(s, sa, max_it, acc));
    probe_next_bytes(33.0 * double(s->ncell));   // markers 1, FP64 pressure 8, faces read 12 and written 12
    MGF_HIP(launch_project(st, s->g, sg, s->markers, s->x_is_64 ? nullptr : s->x,
                           s->x_is_64 ? s->x64 : nullptr, s->u[c], s->v[c], s->w[c], s->us, s->vs,
                           s->ws));
    if (s->div_check)
        MGF_HIP(launch_div_check(st, s->g, s->markers, s->u[c], s->v[c], s->w[c], s->b, s->red.partials,
                                 kDivBlocks, s->div_dev));
    if (!std::isfinite(acc->residual_rel)) return fail(MGF_ERR_NONFINITE, "pressure solve produced NaN/Inf");
    return MGF_OK;
}

// Resolution-derived sizes of a scene.
void shape_scene(mgf_scene s, const mgf_scene_desc* desc) {
    s->desc = *desc;
    s->g.nx = desc->nx; s->g.ny = desc->ny; s->g.nz = desc->nz;
    s->g.ox = desc->origin[0]; s->g.oy = desc->origin[1]; s->g.oz = desc->origin[2];
    s->g.hx = desc->spacing[0]; s->g.hy = desc->spacing[1]; s->g.hz = desc->spacing[2];
    set_recip(s->g);
    s->ncell = size_t(desc->nx) * desc->ny * desc->nz;
    s->nfx = size_t(desc->nx + 1) * desc->ny * desc->nz;
    s->nfy = size_t(desc->nx) * (desc->ny + 1) * desc->nz;
    s->nfz = size_t(desc->nx) * desc->ny * (desc->nz + 1);
    s->nvec = s->ncell;
}

mgf_status alloc_extrapolation(mgf_scene s, size_t nface);

// Every buffer a scene holds from creation on.
mgf_status alloc_scene(mgf_scene s) {
    mgf_status st = MGF_OK;
    auto A = [&](auto** p, size_t n) {
        if (st == MGF_OK) st = dalloc(s, p, n);
    };
    for (int k = 0; k < 2; ++k) {
        A(&s->u[k], s->nfx); A(&s->v[k], s->nfy); A(&s->w[k], s->nfz); A(&s->phi[k], s->ncell);
    }
    A(&s->solid_total, s->ncell); A(&s->solid_mask, s->ncell);
    A(&s->us, s->nfx); A(&s->vs, s->nfy); A(&s->ws, s->nfz);
    A(&s->markers, s->ncell);
    A(&s->b, s->ncell); A(&s->x, s->ncell); A(&s->x64, s->ncell);
    A(&s->r, s->ncell); A(&s->p0, s->ncell); A(&s->p1, s->ncell); A(&s->ap, s->ncell);
    A(&s->codes, s->ncell);
    if (s->desc.precision == MGF_PRECISION_MIXED) A(&s->wet_flat, col_blocks(s->g));
    A(&s->scal, 1);
    A(&s->red.partials, size_t(std::max<uint32_t>(kMaxBlocks, col_blocks(s->g))) * 2);
    A(&s->red.out2, 2);
    A(&s->div_dev, 3);
    if (st == MGF_OK) st = alloc_extrapolation(s, s->nfx + s->nfy + s->nfz);
    return st;
}

// ---------------------------------------------------------------------------
// Extrapolation: distances, counting sort by distance, one gather kernel per
// layer; returns the largest non-empty layer (< layers).
mgf_status alloc_extrapolation(mgf_scene s, size_t nface) {
    const int32_t K = std::min<int32_t>(ext_max_distance(s->g), 65535);
    TRY(dalloc(s, &s->layer, nface));
    TRY(dalloc(s, &s->list, nface));
    TRY(dalloc(s, &s->lmask, nface));
    TRY(dalloc(s, &s->fmask, nface));
    TRY(dalloc(s, &s->counts, ext_counts_size(K)));
    if (s->dry) return MGF_OK;
    if (hipHostMalloc(reinterpret_cast<void**>(&s->host_offs), sizeof(uint32_t) * (size_t(K) + 3)) != hipSuccess) {
        (void)hipGetLastError();
        s->host_offs = nullptr;
        return fail(MGF_ERR_NOMEM, "hipHostMalloc failed");
    }
    return MGF_OK;
}

mgf_status extrapolate(mgf_scene s, float* u, float* v, float* w, int layers, int* used) {
    ExtBufs B{reinterpret_cast<uint16_t*>(s->layer), s->list, s->lmask, s->fmask, s->counts, s->host_offs};
    return extrapolate_run(s->stream, s->g, u, v, w, layers, B, used);
}

mgf_status step_once(mgf_scene s, mgf_step_stats* acc) {
    const bool t = s->timing;
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    if (t) { TRY(sync(s)); t0 = now_ms(); }
    TRY(refresh_boundaries(s));
    TRY(advect(s, true));
    if (t) { TRY(sync(s)); t1 = now_ms(); }
    TRY(pressure(s, acc));
    if (t) { TRY(sync(s)); t2 = now_ms(); }
    int used = 0;
    TRY(extrapolate(s, s->u[s->cur], s->v[s->cur], s->w[s->cur], s->desc.extrapolation_layers, &used));
    acc->extrapolation_layers = used;
    if (t) {
        TRY(sync(s));
        t3 = now_ms();
        acc->ms_advect += t1 - t0;
        acc->ms_pressure += t2 - t1;
        acc->ms_extrapolate += t3 - t2;
        acc->ms_total += t3 - t0;
    }
    acc->steps += 1;
    return MGF_OK;
}

// _onAdvanceTimeStep with a particle advection solver: _computeAdvection
// (GridFluidSolver.cpp:399-416: velocity, colour, fluid domain, all through
// the old velocity, the boundary SDF of the grid's resolution passed so the
// particles are clamped), forces, the projection with particle markers
// (:434-525), extrapolation.
mgf_status step_particles_once(mgf_scene s, mgf_psolver ps, mgf_step_stats* acc) {
    const bool t = s->timing;
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    if (t) { TRY(sync(s)); t0 = now_ms(); }
    TRY(refresh_boundaries(s));
    hipStream_t st = s->stream;
    const int c = s->cur, n = 1 - s->cur;
    const float* in[3] = {s->u[c], s->v[c], s->w[c]};
    float* out[3] = {s->u[n], s->v[n], s->w[n]};
    TRY(psolver_velocity(ps, st, in, in, s->desc.dt, s->desc.rk_order, 1, out));
    if (s->color[0][0]) {
        const float* ic[3] = {s->color[c][0], s->color[c][1], s->color[c][2]};
        float* oc[3] = {s->color[n][0], s->color[n][1], s->color[n][2]};
        TRY(psolver_vector(ps, st, ic, oc, true));
    }
    TRY(psolver_scalar(ps, st, s->phi[c], s->phi[n], true));
    s->cur = n;
    TRY(apply_forces(s));
    if (t) { TRY(sync(s)); t1 = now_ms(); }
    TRY(pressure(s, acc, ps->p));
    if (t) { TRY(sync(s)); t2 = now_ms(); }
    int used = 0;
    TRY(extrapolate(s, s->u[s->cur], s->v[s->cur], s->w[s->cur], s->desc.extrapolation_layers, &used));
    acc->extrapolation_layers = used;
    if (t) {
        TRY(sync(s));
        t3 = now_ms();
        acc->ms_advect += t1 - t0;
        acc->ms_pressure += t2 - t1;
        acc->ms_extrapolate += t3 - t2;
        acc->ms_total += t3 - t0;
    }
    acc->steps += 1;
    return MGF_OK;
}

// ===========================================================================
// Lockstep batched step.  Scenes stepped together run their pressure solves
// as multi-scene launches (k_multi): one launch per kernel of the mixed CG
// and of the V-cycle covers every scene, a block finding its scene and its
// block of that scene's own launch through a map, so a batch of small scenes
// costs the launches of one scene instead of one set per scene.  Every scene
// receives the arguments of its own launches (the rec_* records of the
// launch_* calls cg_mixed and apply_precond make), in its own order, so its
// results are bitwise those of mgf_step.  The stages before and after the
// solve (advection, markers, right-hand side, projection, extrapolation) run
// on the scenes' own streams.

// Lockstep applies to mixed-precision multigrid scenes starting from x = 0.
bool lock_eligible(mgf_scene s) {
    const auto& d = s->desc;
    return d.precision == MGF_PRECISION_MIXED && d.preconditioner == MGF_PRECOND_MULTIGRID && !d.warm_start;
}

// One record of a scene's sequence, keyed by its place: sequences of
// different scenes merge by key (same key, same kernel = one launch).
struct KRec {
    int a, b, c, d;
    MRec r;
};

void keyed(std::vector<KRec>& out, const std::vector<MRec>& recs, int a, int b, int c) {
    for (size_t k = 0; k < recs.size(); ++k) out.push_back(KRec{a, b, c, int(k), recs[k]});
}

// apply_precond's launches for scene s as keyed records; *zout = the buffer
// that ends up holding z.  Keys: (0, level, step) on the way down, (1, 0, 0)
// the coarse cycle, (2, -level, step) on the way up.
mgf_status vcycle_records(mgf_scene s, const float sa[3], std::vector<KRec>& out, const float** zout) {
    const auto& d = s->desc;
    const MgDims d0{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    const int nu = d.mg_sweeps > 0 ? d.mg_sweeps : 3;
    const float omega = d.mg_omega > 0.0f ? d.mg_omega : 0.9f;
    const int L = 1 + int(s->mg.size());
    struct Lv {
        MgDims d; const uint16_t* codes; const float* b; float* x; float* t; const uint32_t* sched; double cells;
    };
    Lv lv[17];
    lv[0] = Lv{d0, s->codes, s->r, s->z, s->zt, s->sched0, s->wet0_cells};
    for (int l = 1; l < L; ++l) {
        auto& m = s->mg[size_t(l - 1)];
        lv[l] = Lv{m.d, m.codes, m.b, m.x, m.t, m.sched, m.march_cells};
    }
    const int skip = s->skip_dry ? 1 : 0;
    const int l0 = s->mg_l0 < L ? s->mg_l0 : L;
    auto fused = [&](int l) { return l + 1 < L && march_on(s, l, lv[l].d, lv[l + 1].d); };
    std::vector<MRec> r;
    for (int l = 0; l < l0; ++l) {
        Lv& v = lv[l];
        const bool coarsest = l == L - 1;
        if (fused(l)) {
            r.clear();
            rec_mg_descend(r, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].codes, const_cast<float*>(lv[l + 1].b),
                           omega, s->scal, v.sched, skip);
            keyed(out, r, 0, l, 0);
            continue;
        }
        const int sweeps = coarsest ? kCoarsestSweeps : nu;
        auto rz_at = [&](int k) { return (l == 0 && coarsest && k == sweeps - 1) ? 1 : 0; };
        int k;
        r.clear();
        if (sweeps >= 2) {
            rec_mg_smooth2_first(r, v.d, v.codes, v.b, v.x, omega, s->scal, s->red, rz_at(1));
            k = 2;
        } else {
            rec_mg_smooth(r, v.d, v.codes, v.b, v.b, v.x, omega, s->scal, s->red, 1, rz_at(0));
            k = 1;
        }
        keyed(out, r, 0, l, 0);
        for (; k < sweeps; ++k) {
            r.clear();
            rec_mg_smooth(r, v.d, v.codes, v.b, v.x, v.t, omega, s->scal, s->red, 0, rz_at(k));
            keyed(out, r, 0, l, k);
            std::swap(v.x, v.t);
        }
        if (!coarsest) {
            r.clear();
            rec_mg_restrict(r, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].codes, const_cast<float*>(lv[l + 1].b),
                            s->scal);
            keyed(out, r, 0, l, 1000);
        }
    }
    if (l0 < L) {
        MgDims dims[kCoarseLevelsMax];
        const uint16_t* codes[kCoarseLevelsMax];
        for (int l = l0; l < L; ++l) {
            dims[l - l0] = lv[l].d;
            codes[l - l0] = lv[l].codes;
        }
        r.clear();
        if (!rec_mg_coarse_cycle(r, dims, codes, L - l0, lv[l0].b, lv[l0].x, nu, kCoarsestSweeps, omega, s->scal))
            return fail(MGF_ERR_INVALID, "coarse cycle levels");
        keyed(out, r, 1, 0, 0);
    }
    for (int l = std::min(L - 2, l0 - 1); l >= 0; --l) {
        Lv& v = lv[l];
        auto rz_at = [&](int k) { return (l == 0 && k == nu - 1) ? 1 : 0; };
        r.clear();
        if (fused(l)) {
            rec_mg_ascend(r, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].x, v.t, omega, s->scal, s->red,
                          l == 0 ? 1 : 0, v.sched, skip);
            keyed(out, r, 2, -l, 0);
            std::swap(v.x, v.t);
            continue;
        }
        rec_mg_smooth_prolong(r, v.d, lv[l + 1].d, v.codes, v.b, v.x, lv[l + 1].x, v.t, omega, s->scal, s->red,
                              rz_at(0));
        keyed(out, r, 2, -l, 0);
        std::swap(v.x, v.t);
        for (int k = 1; k < nu; ++k) {
            r.clear();
            rec_mg_smooth(r, v.d, v.codes, v.b, v.x, v.t, omega, s->scal, s->red, 0, rz_at(k));
            keyed(out, r, 2, -l, k);
            std::swap(v.x, v.t);
        }
    }
    *zout = lv[0].x;
    return MGF_OK;
}

// Allocations and host state the phase records rely on: the flat kernels'
// block flags, the hierarchy and its levels' operator scales (mg_setup sets
// the same scales each solve).
mgf_status lock_prepare(mgf_scene s) {
    if (!s->wet_flat) TRY(dalloc(s, &s->wet_flat, col_blocks(s->g)));
    TRY(ensure_precond(s));
    float sa[3], sg[3];
    scales(s, sa, sg);
    MgDims f{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    for (auto& L : s->mg) {
        L.d.sx = f.nx > 1 ? f.sx * 0.25f : f.sx;
        L.d.sy = f.ny > 1 ? f.sy * 0.25f : f.sy;
        L.d.sz = f.nz > 1 ? f.sz * 0.25f : f.sz;
        f = L.d;
    }
    return MGF_OK;
}

// The launches of a step before the CG sweeps for a scene whose current
// buffers are `c` (advection with forces, markers, right-hand side, stencil
// codes, the hierarchy's markers and codes, x64 = 0 and ||b||^2: advect,
// pressure() and cg_mixed's prologue), and after them (projection with the
// FP64 pressure, the extrapolation's distance transform and sort, its layer
// offsets and largest layer copied to `offs_out`), as keyed records; the
// layers themselves are launched from those offsets.
mgf_status phase_records(mgf_scene s, int c, std::vector<KRec>& pre, std::vector<KRec>& post, uint32_t* offs_out) {
    const int n = 1 - c;
    const float dt = s->desc.dt;
    float sa[3], sg[3];
    scales(s, sa, sg);
    float g[3], f[3];
    for (int k = 0; k < 3; ++k) {
        g[k] = s->desc.gravity[k] * dt;
        f[k] = s->desc.external_force[k] * dt;
    }
    auto on = [](const float* a) {
        return !(std::fabs(a[0]) < MGF_FLT_EPSILON && std::fabs(a[1]) < MGF_FLT_EPSILON &&
                 std::fabs(a[2]) < MGF_FLT_EPSILON);
    };
    std::vector<MRec> r;
    const Vel V = vel_of(s->u[c], s->v[c], s->w[c]);
    rec_advect_fc(r, s->g, V, V, dt, s->desc.rk_order, s->u[n], s->v[n], s->w[n], on(g), g, on(f), f);
    keyed(pre, r, 0, 0, 0);
    r.clear();
    if (s->color[0][0]) {
        const float* ic[3] = {s->color[c][0], s->color[c][1], s->color[c][2]};
        float* oc[3] = {s->color[n][0], s->color[n][1], s->color[n][2]};
        rec_advect_cc4(r, s->g, s->phi[c], ic, V, dt, s->desc.rk_order, s->phi[n], oc);
    } else {
        rec_advect_cc(r, s->g, s->phi[c], V, dt, s->desc.rk_order, s->phi[n]);
    }
    keyed(pre, r, 0, 1, 0);
    r.clear();
    rec_markers(r, s->g, s->solid_mask, s->phi[n], s->markers);
    keyed(pre, r, 0, 2, 0);
    r.clear();
    rec_rhs(r, s->g, vel_of(s->u[n], s->v[n], s->w[n]), vel_of(s->us, s->vs, s->ws), s->markers, s->b);
    keyed(pre, r, 0, 3, 0);
    r.clear();
    rec_codes(r, s->g.nx, s->g.ny, s->g.nz, s->markers, s->codes);
    keyed(pre, r, 0, 4, 0);
    r.clear();
    rec_flat_wet(r, s->g, s->codes, s->wet_flat);
    keyed(pre, r, 0, 5, 0);
    // mg_setup
    MgDims fd{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    const uint8_t* mf = s->markers;
    for (size_t l = 0; l < s->mg.size(); ++l) {
        auto& L = s->mg[l];
        r.clear();
        rec_mg_coarsen(r, fd, L.d, mf, L.markers, 0);
        rec_codes(r, L.d.nx, L.d.ny, L.d.nz, L.markers, L.codes);
        keyed(pre, r, 1, int(l), 0);
        fd = L.d;
        mf = L.markers;
    }
    fd = MgDims{s->g.nx, s->g.ny, s->g.nz, sa[0], sa[1], sa[2]};
    const uint16_t* cf = s->codes;
    for (size_t l = 0; l < s->mg.size(); ++l) {
        const MgDims& cd = s->mg[l].d;
        uint32_t* sched = l == 0 ? s->sched0 : s->mg[l - 1].sched;
        const size_t sched_n = l == 0 ? s->sched0_n : s->mg[l - 1].sched_n;
        if (sched && sched_n >= mg_march_sched_words(fd, cd)) {
            r.clear();
            rec_mg_march_wet(r, fd, cd, cf, sched);
            rec_mg_march_plan(r, fd, cd, sched);
            keyed(pre, r, 2, int(l), 0);
        }
        fd = cd;
        cf = s->mg[l].codes;
    }
    r.clear();
    rec_refine_true(r, s->g, sa, s->codes, nullptr, s->x64, s->b, nullptr, 0, s->red, s->red.out2);
    keyed(pre, r, 3, 0, 0);
    // after the solve
    r.clear();
    rec_project(r, s->g, sg, s->markers, nullptr, s->x64, s->u[n], s->v[n], s->w[n], s->us, s->vs, s->ws);
    keyed(post, r, 0, 0, 0);
    const int32_t K = ext_max_distance(s->g);
    const int32_t cap = int32_t(std::min<int64_t>(int64_t(s->desc.extrapolation_layers) - 1, K));
    if (cap > 0) {
        if (K >= 65535)
            return fail(MGF_ERR_UNSUPPORTED, "extrapolation over more than one layer needs nx + ny + nz + 3 < 65535 "
                                             "(16-bit face distances)");
        r.clear();
        const uint32_t* maxd = nullptr;
        rec_ext_device(r, s->g, s->u[n], s->v[n], s->w[n], reinterpret_cast<uint16_t*>(s->layer), cap, K, s->counts,
                       s->list, s->lmask, s->fmask, &maxd, false);
        keyed(post, r, 1, 0, 0);
        r.clear();
        rec_copy_offs(r, s->counts, K, cap, offs_out);
        keyed(post, r, 2, 0, 0);
    }
    return MGF_OK;
}

// A sequence of multi-scene launches merged from the scenes' keyed records.
struct LSlot {
    int32_t kind;
    uint32_t nblocks, shmem;
    size_t tab, map;   // byte offsets into the plan's device buffer
};
using LSeq = std::vector<LSlot>;

constexpr int kActiveRing = 8;

struct LockPlan {
    mgf_context ctx = nullptr;
    std::vector<mgf_scene> sc;
    hipStream_t stream = nullptr;
    unsigned char* dev = nullptr;
    uint8_t* active_dev = nullptr;
    CgScalars* gsc_dev = nullptr;
    double* gout_dev = nullptr;
    unsigned char* host = nullptr;   // pinned: gathered scalars and sums, then kActiveRing active-flag slots
    uint8_t* active_host = nullptr;
    size_t active_stride = 0;
    int ring = 0;                    // active-flag slots used since the last synchronisation
    CgScalars* gsc_host = nullptr;
    double* gout_host = nullptr;
    LSeq start_b, start_r, cg, vcycle, refine1, refine2, gather;
    std::vector<const float*> zv;
    std::vector<unsigned char> blob;   // host copy of the tables and maps
    hipGraphExec_t iter = nullptr;     // one CG iteration (cg + vcycle), captured
    ~LockPlan() {
        if (iter) (void)hipGraphExecDestroy(iter);
        if (ctx) {
            pool_put(ctx, dev);
            pool_put(ctx, host);
        }
    }
};

// Merges per-scene keyed records into launches and lays out their tables and
// maps from byte offset `at` on, at 256-byte boundaries, advancing `at` past
// them.  With blob null only the offsets are computed (the size to
// allocate); otherwise blob holds the final size, zero where the tables go.
LSeq merge_seq(const std::vector<std::vector<KRec>>& per, unsigned char* blob, size_t& at) {
    struct Key {
        int a, b, c, d, kind;
        bool operator<(const Key& o) const {
            if (a != o.a) return a < o.a;
            if (b != o.b) return b < o.b;
            if (c != o.c) return c < o.c;
            if (d != o.d) return d < o.d;
            return kind < o.kind;
        }
    };
    std::vector<std::pair<Key, std::pair<int, const MRec*>>> all;
    for (size_t i = 0; i < per.size(); ++i)
        for (const KRec& k : per[i]) all.push_back({Key{k.a, k.b, k.c, k.d, k.r.kind}, {int(i), &k.r}});
    std::stable_sort(all.begin(), all.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    auto align = [&]() { at = (at + 255) / 256 * 256; };
    LSeq seq;
    for (size_t i = 0; i < all.size();) {
        size_t j = i;
        while (j < all.size() && !(all[i].first < all[j].first) && !(all[j].first < all[i].first)) ++j;
        LSlot sl{all[i].first.kind, 0, 0, 0, 0};
        align();
        // one entry per scene of the batch, at the scene's index (the map
        // names the scene); scenes without this launch leave theirs zero
        sl.tab = at;
        const size_t bytes = all[i].second.second->bytes;
        at += bytes * per.size();
        std::vector<uint8_t> seen(per.size(), 0);
        for (size_t q = i; q < j; ++q) {
            const MRec* r = all[q].second.second;
            const size_t scene = size_t(all[q].second.first);
            // a scene has at most one record per key (its entry would be overwritten)
            if (seen[scene]++ || r->bytes != bytes) {
                seq.clear();
                return seq;
            }
            if (blob) std::memcpy(blob + sl.tab + scene * bytes, r->args, bytes);
            sl.shmem = std::max(sl.shmem, r->shmem);
            sl.nblocks += r->blocks;
        }
        align();
        sl.map = at;
        at += 4 * size_t(sl.nblocks);
        if (blob) {
            unsigned char* mp = blob + sl.map;
            for (size_t q = i; q < j; ++q) {
                const uint32_t scene = uint32_t(all[q].second.first) << kMapShift;
                const uint32_t nb = all[q].second.second->blocks;
                for (uint32_t b = 0; b < nb; ++b, mp += 4) {
                    const uint32_t e = scene | b;
                    std::memcpy(mp, &e, 4);
                }
            }
        }
        seq.push_back(sl);
        i = j;
    }
    return seq;
}

mgf_status run_seq_at(hipStream_t st, unsigned char* base, const LSeq& seq, const uint8_t* active) {
    for (const LSlot& sl : seq)
        MGF_HIP(launch_multi(st, sl.kind, base + sl.tab, reinterpret_cast<const uint32_t*>(base + sl.map), sl.nblocks,
                             sl.shmem, active));
    return MGF_OK;
}

mgf_status run_seq(LockPlan& P, const LSeq& seq) { return run_seq_at(P.stream, P.dev, seq, P.active_dev); }

// The launches before and after the solves for one assignment of current
// buffers (each scene's `cur`): built once per assignment and kept (a batch
// alternates between two).
struct PhasePlan {
    mgf_context ctx = nullptr;
    LSeq pre, post;
    unsigned char* dev = nullptr;   // the packed layer offsets, the layer table, then tables and maps
    std::vector<size_t> offs_at;    // per scene: its offsets' first word in the packed area (cap + 3 words)
    std::vector<int32_t> cap;       // per scene: the layer cap (0: no extrapolation)
    size_t offs_words = 0;
    size_t layer_tab = 0;           // byte offset of the per-scene k_multi_layer arguments
    uint32_t* offs_host = nullptr;  // pinned copy of the packed offsets
    PhasePlan() = default;
    PhasePlan(const PhasePlan&) = delete;
    PhasePlan& operator=(const PhasePlan&) = delete;
    ~PhasePlan() {
        if (ctx) {
            pool_put(ctx, dev);
            pool_put(ctx, offs_host);
        }
    }
};

mgf_status build_phase(const std::vector<mgf_scene>& sc, const std::vector<uint8_t>& cur, hipStream_t st,
                       PhasePlan& Q) {
    const size_t n = sc.size();
    Q.offs_at.assign(n, 0);
    Q.cap.assign(n, 0);
    Q.offs_words = 0;
    for (size_t i = 0; i < n; ++i) {
        const int32_t K = ext_max_distance(sc[i]->g);
        Q.cap[i] = int32_t(std::max<int64_t>(0, std::min<int64_t>(int64_t(sc[i]->desc.extrapolation_layers) - 1, K)));
        Q.offs_at[i] = Q.offs_words;
        if (Q.cap[i] > 0) Q.offs_words += size_t(Q.cap[i]) + 3;
    }
    Q.layer_tab = (Q.offs_words * sizeof(uint32_t) + 255) / 256 * 256;
    const size_t head = Q.layer_tab + (n * layer_arg_bytes() + 255) / 256 * 256;
    std::vector<std::vector<KRec>> pre(n), post(n);
    // the layout first (records with null addresses), then allocate and fill
    size_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        uint32_t* packed = reinterpret_cast<uint32_t*>(pass == 0 ? nullptr : Q.dev);
        for (size_t i = 0; i < n; ++i) {
            pre[i].clear();
            post[i].clear();
            TRY(phase_records(sc[i], cur[i], pre[i], post[i], packed + Q.offs_at[i]));
        }
        std::vector<unsigned char> blob(pass == 0 ? 0 : total, 0);
        unsigned char* bp = pass == 0 ? nullptr : blob.data();
        for (size_t i = 0; bp && i < n; ++i) {
            mgf_scene s = sc[i];
            const int c1 = 1 - cur[i];
            layer_args(bp + Q.layer_tab + i * layer_arg_bytes(), s->g, s->u[c1], s->v[c1], s->w[c1], s->list,
                       s->lmask, s->counts, ext_max_distance(s->g));
        }
        size_t at = head;
        Q.pre = merge_seq(pre, bp, at);
        Q.post = merge_seq(post, bp, at);
        if (Q.pre.empty() || Q.post.empty()) return fail(MGF_ERR_INVALID, "lockstep phases: a scene has two records under one key");
        if (pass == 0) {
            total = at;
            Q.ctx = sc[0]->ctx;
            Q.dev = static_cast<unsigned char*>(pool_get(Q.ctx, total, false));
            if (!Q.dev) return fail(MGF_ERR_NOMEM, "lockstep phases: hipMalloc failed");
            Q.offs_host = static_cast<uint32_t*>(pool_get(Q.ctx, std::max<size_t>(Q.offs_words, 1) * 4, true));
            if (!Q.offs_host) return fail(MGF_ERR_NOMEM, "lockstep phases: hipHostMalloc failed");
            continue;
        }
        if (at != total) return fail(MGF_ERR_INVALID, "lockstep phases: the layout changed between its passes");
        MGF_HIP(hipMemcpyAsync(Q.dev, bp, total, hipMemcpyHostToDevice, st));
        MGF_HIP(hipStreamSynchronize(st));
    }
    return MGF_OK;
}

// Builds the plan for scenes whose hierarchies exist (after lock_prepare).
mgf_status build_plan(LockPlan& P) {
    const size_t n = P.sc.size();
    if (n >= (size_t(1) << (32 - kMapShift))) return fail(MGF_ERR_INVALID, "too many scenes for one lockstep batch");
    std::vector<std::vector<KRec>> start_b(n), start_r(n), cg(n), vc(n), ref1(n), ref2(n), gat(n);
    P.zv.assign(n, nullptr);
    // the gather outputs live in the plan's buffer: their offsets are known
    // before the device buffer exists, so records hold device addresses
    // computed from a base fixed below
    const size_t gsc_off = 0, gout_off = (n * sizeof(CgScalars) + 255) / 256 * 256,
                 act_off = gout_off + (n * 2 * sizeof(double) + 255) / 256 * 256,
                 act_stride = (n + 255) / 256 * 256, head = act_off + act_stride;
    // size the tables first with null addresses, then allocate and fill
    size_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        unsigned char* base = pass == 0 ? nullptr : P.dev;
        CgScalars* gsc = reinterpret_cast<CgScalars*>(base + gsc_off);
        double* gout = reinterpret_cast<double*>(base + gout_off);
        for (size_t i = 0; i < n; ++i) {
            mgf_scene s = P.sc[i];
            float sa[3], sg[3];
            scales(s, sa, sg);
            const double inner = s->desc.inner_tolerance > 0 ? s->desc.inner_tolerance : 1e-3;
            const uint8_t* wet_flat = s->skip_dry ? s->wet_flat : nullptr;
            std::vector<MRec> r;
            start_b[i].clear(); start_r[i].clear(); cg[i].clear(); vc[i].clear();
            ref1[i].clear(); ref2[i].clear(); gat[i].clear();
            // the overall target of each sweep's start is set per step
            // (set_start_targets)
            rec_cg_start(r, s->g, s->b, s->r, s->x, s->scal, s->red, inner, 0.0);
            keyed(start_b[i], r, 0, 0, 0);
            r.clear();
            rec_cg_start(r, s->g, s->r, s->r, s->x, s->scal, s->red, inner, 0.0);
            keyed(start_r[i], r, 0, 0, 0);
            TRY(vcycle_records(s, sa, vc[i], &P.zv[i]));
            r.clear();
            rec_col_p_ap(r, s->g, sa, s->codes, P.zv[i], s->p0, s->p1, s->scal, s->red, wet_flat);
            keyed(cg[i], r, 0, 0, 0);
            r.clear();
            rec_col_update(r, s->g, s->p0, s->p1, s->x, s->r, s->scal, s->red, 0, wet_flat);
            keyed(cg[i], r, 0, 0, 1);
            r.clear();
            rec_refine_true(r, s->g, sa, s->codes, s->x, s->x64, s->b, s->r, 1, s->red, s->red.out2);
            keyed(ref1[i], r, 0, 0, 0);
            r.clear();
            rec_refine_true(r, s->g, sa, s->codes, s->x, s->x64, s->b, s->r, 2, s->red, s->red.out2);
            keyed(ref2[i], r, 0, 0, 0);
            r.clear();
            rec_gather(r, s->scal, s->red.out2, gsc + i, gout + 2 * i);
            keyed(gat[i], r, 0, 0, 0);
        }
        std::vector<unsigned char>& blob = P.blob;
        blob.assign(pass == 0 ? 0 : total, 0);
        unsigned char* bp = pass == 0 ? nullptr : blob.data();
        size_t at = head;
        P.start_b = merge_seq(start_b, bp, at);
        P.start_r = merge_seq(start_r, bp, at);
        P.cg = merge_seq(cg, bp, at);
        P.vcycle = merge_seq(vc, bp, at);
        P.refine1 = merge_seq(ref1, bp, at);
        P.refine2 = merge_seq(ref2, bp, at);
        P.gather = merge_seq(gat, bp, at);
        for (const LSeq* q : {&P.start_b, &P.start_r, &P.cg, &P.vcycle, &P.refine1, &P.refine2, &P.gather})
            if (q->empty()) return fail(MGF_ERR_INVALID, "lockstep plan: a scene has two records under one key");
        if (pass == 1 && at != total) return fail(MGF_ERR_INVALID, "lockstep plan: the layout changed between its passes");
        total = at;
        if (pass == 0) {
            P.dev = static_cast<unsigned char*>(pool_get(P.ctx, total, false));
            if (!P.dev) return fail(MGF_ERR_NOMEM, "lockstep plan: hipMalloc failed");
            const size_t hb = act_off + kActiveRing * act_stride;
            P.host = static_cast<unsigned char*>(pool_get(P.ctx, hb, true));
            if (!P.host) return fail(MGF_ERR_NOMEM, "lockstep plan: hipHostMalloc failed");
            continue;
        }
        MGF_HIP(hipMemcpyAsync(P.dev, blob.data(), blob.size(), hipMemcpyHostToDevice, P.stream));
        MGF_HIP(hipStreamSynchronize(P.stream));
    }
    P.gsc_dev = reinterpret_cast<CgScalars*>(P.dev + gsc_off);
    P.gout_dev = reinterpret_cast<double*>(P.dev + gout_off);
    P.active_dev = P.dev + act_off;
    P.gsc_host = reinterpret_cast<CgScalars*>(P.host + gsc_off);
    P.gout_host = reinterpret_cast<double*>(P.host + gout_off);
    P.active_host = P.host + act_off;
    P.active_stride = act_stride;
    P.ring = 0;
    return MGF_OK;
}

// Captures one CG iteration of every scene (K_I, K_II with their finishing
// kernels, the V-cycle) into a graph: its launches and tables do not change
// within a batch call, so a chunk of k iterations is k launches of the graph.
mgf_status capture_iteration(LockPlan& P) {
    hipGraph_t g = nullptr;
    MGF_HIP(hipStreamBeginCapture(P.stream, hipStreamCaptureModeRelaxed));
    mgf_status st = run_seq(P, P.cg);
    if (st == MGF_OK) st = run_seq(P, P.vcycle);
    const hipError_t e = hipStreamEndCapture(P.stream, &g);
    if (st != MGF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MGF_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&P.iter, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        (void)hipGetLastError();
        P.iter = nullptr;
        return fail(MGF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    return MGF_OK;
}

// Gathers every scene's scalars and refinement sums to the host.
mgf_status gather(LockPlan& P) {
    const size_t n = P.sc.size();
    // the gather runs for every scene, active or not
    MGF_HIP(hipMemsetAsync(P.active_dev, 1, n, P.stream));
    TRY(run_seq(P, P.gather));
    MGF_HIP(hipMemcpyAsync(P.host, P.dev, reinterpret_cast<unsigned char*>(P.active_dev) - P.dev, hipMemcpyDeviceToHost,
                           P.stream));
    MGF_HIP(hipStreamSynchronize(P.stream));
    P.ring = 0;
    return MGF_OK;
}

// Copies the active flags in through a pinned slot that no copy still in
// flight reads (the slots are reused after the next synchronisation).
mgf_status set_active(LockPlan& P, const std::vector<uint8_t>& act) {
    if (P.ring >= kActiveRing) {
        MGF_HIP(hipStreamSynchronize(P.stream));
        P.ring = 0;
    }
    uint8_t* slot = P.active_host + size_t(P.ring++) * P.active_stride;
    std::memcpy(slot, act.data(), act.size());
    MGF_HIP(hipMemcpyAsync(P.active_dev, slot, act.size(), hipMemcpyHostToDevice, P.stream));
    return MGF_OK;
}

// Sets each scene's overall stop target in the finishing records of the
// sweep starts (the only argument that changes between steps).
mgf_status set_start_targets(LockPlan& P, const std::vector<double>& target2) {
    for (LSeq* q : {&P.start_b, &P.start_r})
        for (const LSlot& sl : *q) {
            if (sl.kind != kMkFinish) continue;
            const size_t bytes = finish_arg_bytes();
            for (size_t i = 0; i < target2.size(); ++i) finish_set_target(P.blob.data() + sl.tab + i * bytes, target2[i]);
            MGF_HIP(hipMemcpyAsync(P.dev + sl.tab, P.blob.data() + sl.tab, bytes * target2.size(), hipMemcpyHostToDevice,
                                   P.stream));
        }
    return MGF_OK;
}

// The sweeps of cg_mixed for every scene at once (after the pre-solve phase
// scene's stream and a wait for them on the plan's stream).
mgf_status lock_solve(LockPlan& P, std::vector<mgf_step_stats>& acc) {
    const size_t n = P.sc.size();
    TRY(gather(P));
    struct St {
        double b2 = 0, target2 = 0, rr_true = 0, inner = 1e-3, log_rho = 0, rr0 = 0, sweep_target = 0, rr_now = 0;
        int64_t it_total = 0, max_it = 0;
        int sweeps = 0, max_sweeps = 8, done_before = 0, chunk = 16, first_chunk = 4;
        bool conv = false, stop = false, running = false, nan = false;
        CgScalars hs{};
    };
    std::vector<St> S(n);
    for (size_t i = 0; i < n; ++i) {
        mgf_scene s = P.sc[i];
        const auto& d = s->desc;
        St& t = S[i];
        t.inner = d.inner_tolerance > 0 ? d.inner_tolerance : 1e-3;
        t.max_sweeps = d.max_refinements > 0 ? d.max_refinements : 8;
        t.max_it = d.max_iterations > 0 ? d.max_iterations : int64_t(s->ncell);
        t.b2 = P.gout_host[2 * i + 1];
        if (d.tol_mode == MGF_TOL_ABS_SQRT_NORM) {
            const double q = d.tolerance * d.tolerance;
            t.target2 = q * q;
        } else {
            t.target2 = d.tolerance * d.tolerance * t.b2;
        }
        t.rr_true = t.b2;
        t.conv = t.rr_true <= t.target2 || t.b2 == 0.0;
        s->x_is_64 = true;
    }
    {
        std::vector<double> tg(n);
        for (size_t i = 0; i < n; ++i) tg[i] = S[i].target2;
        TRY(set_start_targets(P, tg));
    }
    auto predict = [](const St& t) {
        if (!(t.log_rho < 0.0) || !(t.rr_now > t.sweep_target) || !(t.sweep_target > 0.0)) return t.first_chunk;
        const double k = std::floor(std::log(t.sweep_target / t.rr_now) / t.log_rho + 0.5);
        return int(std::max(1.0, std::min(double(t.chunk), k)));
    };
    std::vector<uint8_t> act(n, 0);
    for (;;) {
        bool any = false;
        for (size_t i = 0; i < n; ++i) {
            St& t = S[i];
            act[i] = (!t.conv && !t.stop && t.sweeps < t.max_sweeps && t.it_total < t.max_it) ? 1 : 0;
            any = any || act[i];
        }
        if (!any) break;
        TRY(set_active(P, act));
        // each scene's first sweep starts from b, later ones from the refined residual
        std::vector<uint8_t> from_b(n, 0), from_r(n, 0);
        for (size_t i = 0; i < n; ++i) {
            from_b[i] = act[i] && S[i].sweeps == 0;
            from_r[i] = act[i] && S[i].sweeps > 0;
        }
        if (std::count(from_b.begin(), from_b.end(), 1)) {
            TRY(set_active(P, from_b));
            TRY(run_seq(P, P.start_b));
        }
        if (std::count(from_r.begin(), from_r.end(), 1)) {
            TRY(set_active(P, from_r));
            TRY(run_seq(P, P.start_r));
        }
        TRY(set_active(P, act));
        TRY(run_seq(P, P.vcycle));
        for (size_t i = 0; i < n; ++i) {
            if (!act[i]) continue;
            St& t = S[i];
            t.rr0 = t.rr_true;
            t.sweep_target = std::max(t.inner * t.inner * t.rr0, t.target2);
            t.rr_now = t.rr0;
            t.done_before = 0;
            t.running = true;
        }
        for (;;) {
            int k = 0;
            int64_t left = std::numeric_limits<int64_t>::max();
            for (size_t i = 0; i < n; ++i) {
                if (!act[i] || !S[i].running) continue;
                k = std::max(k, predict(S[i]));
                left = std::min(left, std::max<int64_t>(S[i].max_it - S[i].it_total - S[i].done_before, 1));
            }
            if (k == 0) break;
            k = int(std::min<int64_t>(k, left));
            for (int q = 0; q < k; ++q) MGF_HIP(hipGraphLaunch(P.iter, P.stream));
            TRY(gather(P));
            TRY(set_active(P, act));
            for (size_t i = 0; i < n; ++i) {
                if (!act[i] || !S[i].running) continue;
                St& t = S[i];
                t.hs = P.gsc_host[i];
                if (t.hs.iters > 0 && t.hs.rr > 0.0 && t.rr0 > 0.0 && t.hs.rr < t.rr0)
                    t.log_rho = std::log(t.hs.rr / t.rr0) / double(t.hs.iters);
                t.rr_now = t.hs.rr;
                t.done_before = t.hs.iters;
                if (t.hs.done || t.it_total + t.hs.iters >= t.max_it) t.running = false;
            }
        }
        // the refinement of every scene that swept
        std::vector<uint8_t> r1(n, 0), r2(n, 0);
        for (size_t i = 0; i < n; ++i) {
            if (!act[i]) continue;
            St& t = S[i];
            t.it_total += t.hs.iters;
            if (t.hs.done == 2) {
                t.rr_true = std::nan("");
                t.nan = true;
                t.stop = true;
                continue;
            }
            (t.sweeps > 0 ? r2 : r1)[i] = 1;
        }
        if (std::count(r1.begin(), r1.end(), 1)) {
            TRY(set_active(P, r1));
            TRY(run_seq(P, P.refine1));
        }
        if (std::count(r2.begin(), r2.end(), 1)) {
            TRY(set_active(P, r2));
            TRY(run_seq(P, P.refine2));
        }
        TRY(gather(P));
        for (size_t i = 0; i < n; ++i) {
            if (!(r1[i] || r2[i])) continue;
            St& t = S[i];
            t.rr_true = P.gout_host[2 * i];
            ++t.sweeps;
            if (t.rr_true <= t.target2) t.conv = true;
            if (!std::isfinite(t.rr_true)) t.stop = true;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        mgf_scene s = P.sc[i];
        St& t = S[i];
        s->x64_valid = t.conv && std::isfinite(t.rr_true);
        mgf_step_stats& a = acc[i];
        a.cg_iterations += t.it_total;
        a.refinement_sweeps += t.sweeps;
        a.converged = a.converged && t.conv;
        a.true_converged = a.true_converged && t.conv;
        a.residual_rel = t.b2 > 0 ? std::sqrt(t.rr_true / t.b2) : std::sqrt(t.rr_true);
        a.residual_recurrence = std::sqrt(t.hs.rr);
        a.b_norm = std::sqrt(t.b2);
    }
    return MGF_OK;
}

void stats_reset(mgf_step_stats* st) {
    std::memset(st, 0, sizeof(*st));
    st->converged = 1;
    st->true_converged = 1;
    st->div_residual = -1.0;
}

// The divergence check's running maximum: cleared before the solves of one
// entry-point call, read after its final synchronisation.
mgf_status div_begin(mgf_scene s) {
    if (!s->div_check) return MGF_OK;
    MGF_HIP(launch_fill64(s->stream, s->div_dev, 1, -1.0));
    return MGF_OK;
}

mgf_status div_end(mgf_scene s, mgf_step_stats* acc) {
    if (!s->div_check) return MGF_OK;
    double h = -1.0;
    MGF_HIP(hipMemcpyAsync(&h, s->div_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    MGF_HIP(hipStreamSynchronize(s->stream));
    acc->div_residual = h;
    return MGF_OK;
}

// ---------------------------------------------------------------------------
// Scratch of mgf_op_extrapolate_fc_noalloc: one file-scope device array, so
// that the entry point allocates nothing.  Every device the library runs on
// holds its own copy until the process ends; its address is looked up once
// per device at mgf_init (hipGetSymbolAddress with that device current).
// Layout (each part 256-byte aligned): face distances uint16[kNaFaces],
// sorted list uint32[kNaFaces], list masks uint8[kNaFaces], face masks
// uint8[kNaFaces], sort counters uint32[kNaCountWords].
constexpr size_t kNaFaces = size_t(1) << 22;   // faces of all three components
constexpr int32_t kNaMaxK = 4093;              // nx + ny + nz + 3 with more than one layer
constexpr size_t kNaCountWords = (3 * 32 + 1) * (size_t(kNaMaxK) + 2) + 1;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaDistOff = 0;
constexpr size_t kNaListOff = kNaDistOff + na_align(2 * kNaFaces);
constexpr size_t kNaLmaskOff = kNaListOff + na_align(4 * kNaFaces);
constexpr size_t kNaFmaskOff = kNaLmaskOff + na_align(kNaFaces);
constexpr size_t kNaCountOff = kNaFmaskOff + na_align(kNaFaces);
constexpr size_t kNaBytes = kNaCountOff + na_align(4 * kNaCountWords);
static_assert(kNaBytes <= kNoAllocScratchBytes, "the scratch array holds the layout");

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};
std::string g_na_err[kNaMaxDevices];

// With `device` current: the device address of this device's scratch.
void na_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_na_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_na_ptr[device] = p;
}

void* na_ptr(int device, std::string* err) {
    if (device < 0 || device >= kNaMaxDevices) {
        *err = "device index beyond the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) *err = g_na_err[device].empty() ? "scratch not set up" : g_na_err[device];
    return g_na_ptr[device];
}

}  // namespace

namespace mgf {

mgf_status scene_bytes_for(mgf_context ctx, const mgf_scene_desc* desc, int64_t* out) {
    if (!ctx || !out) return fail(MGF_ERR_INVALID, "null argument");
    TRY(validate_desc(desc));
    mgf_scene_s dry;
    dry.ctx = ctx;
    dry.dry = true;
    shape_scene(&dry, desc);
    TRY(alloc_scene(&dry));
    if (desc->precision == MGF_PRECISION_MIXED && desc->preconditioner != MGF_PRECOND_IDENTITY)
        TRY(ensure_precond(&dry));
    *out = dry.device_bytes;
    return MGF_OK;
}

mgf_status extrapolate_run(hipStream_t st, const GridDev& g, float* u, float* v, float* w, int layers,
                           const ExtBufs& B, int* used, bool device_only) {
    *used = 0;
    const int32_t K = ext_max_distance(g);
    const int32_t cap = int32_t(std::min<int64_t>(int64_t(layers) - 1, K));
    if (cap <= 0) return MGF_OK;   // layer 0 only: the reference's fill loop is empty
    if (K >= 65535)
        return fail(MGF_ERR_UNSUPPORTED, "extrapolation over more than one layer needs nx + ny + nz + 3 < 65535 "
                                         "(16-bit face distances)");
    uint16_t* dist = B.dist;
    MGF_HIP(launch_ext_distance(st, g, u, v, w, dist));
    const uint32_t* offs = nullptr;
    const uint32_t* maxd = nullptr;
    MGF_HIP(launch_ext_sort(st, g, dist, cap, K, B.counts, B.list, B.lmask, B.fmask, &offs, &maxd));
    if (device_only) {
        probe_next_bytes(0.0);
        MGF_HIP(launch_ext_tail(st, g, u, v, w, B.list, B.lmask, offs, 1, 0, maxd));
        MGF_HIP(hipMemcpyAsync(B.host_offs, maxd, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        *used = -1;
        return MGF_OK;
    }
    // offsets of layers 0..cap+1, then the largest distance present
    MGF_HIP(hipMemcpyAsync(B.host_offs, offs, sizeof(uint32_t) * (size_t(cap) + 2), hipMemcpyDeviceToHost, st));
    MGF_HIP(hipMemcpyAsync(B.host_offs + cap + 2, maxd, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MGF_HIP(hipStreamSynchronize(st));
    const int32_t top = int32_t(B.host_offs[cap + 2]);
    // Small layers are launch-bound: two of them share a launch (the second
    // recomputes its previous-layer neighbours, several loads per face).
    // Large layers are bound by their gathers, which the pairing would
    // multiply, so they run one per launch.
    constexpr uint32_t kPairMax = 65536;   // faces of two layers launched together
    // From the first layer after which no layer exceeds kTailMax entries on,
    // the layers run in one work-group (k_ext_tail): a launch per layer
    // costs ~8 us, a small layer ~2-4 us there.
    constexpr uint32_t kTailMax = 1024;
    int32_t tail = top + 1;
    while (tail > 1 && B.host_offs[tail] - B.host_offs[tail - 1] <= kTailMax) --tail;
    for (int32_t L = 1; L < tail;) {
        const uint32_t b = B.host_offs[L], m = B.host_offs[L + 1];
        const uint32_t e = L + 1 < tail ? B.host_offs[L + 2] : m;
        const bool pair = L + 1 < tail && e - b <= kPairMax;
        // per entry: list 4 + mask 1 + write 4 + the values of the previous-layer neighbours (one on average)
        probe_next_bytes(13.0 * double((pair ? e : m) - b));
        if (pair)
            MGF_HIP(launch_ext_layer2(st, g, u, v, w, B.list, B.lmask, B.fmask, b, m - b, e - m));
        else
            MGF_HIP(launch_ext_layer(st, g, u, v, w, B.list, B.lmask, b, m - b));
        L += pair ? 2 : 1;
    }
    if (tail <= top) {
        probe_next_bytes(13.0 * double(B.host_offs[top + 1] - B.host_offs[tail]));
        MGF_HIP(launch_ext_tail(st, g, u, v, w, B.list, B.lmask, offs, tail, top));
    }
    *used = top;
    return MGF_OK;
}

void* pool_get(mgf_context ctx, size_t bytes, bool pinned) {
    std::lock_guard<std::mutex> lk(ctx->pool.mu);
    CtxPool::Buf* best = nullptr;
    for (auto& b : ctx->pool.bufs)
        if (!b.used && b.pinned == pinned && b.bytes >= bytes && (!best || b.bytes < best->bytes)) best = &b;
    if (best) {
        best->used = true;
        return best->p;
    }
    void* p = nullptr;
    const size_t n = std::max<size_t>(bytes, 256);
    const hipError_t e = pinned ? hipHostMalloc(&p, n) : hipMalloc(&p, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    ctx->pool.bufs.push_back(CtxPool::Buf{p, n, pinned, true});
    return p;
}

void pool_put(mgf_context ctx, void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(ctx->pool.mu);
    for (auto& b : ctx->pool.bufs)
        if (b.p == p) b.used = false;
}

hipStream_t pool_stream(mgf_context ctx) {
    {
        std::lock_guard<std::mutex> lk(ctx->pool.mu);
        if (!ctx->pool.streams.empty()) {
            hipStream_t s = ctx->pool.streams.back();
            ctx->pool.streams.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return s;
}

void pool_stream_put(mgf_context ctx, hipStream_t s) {
    if (!s) return;
    std::lock_guard<std::mutex> lk(ctx->pool.mu);
    ctx->pool.streams.push_back(s);
}

void pool_release(mgf_context ctx) {
    std::lock_guard<std::mutex> lk(ctx->pool.mu);
    for (auto& b : ctx->pool.bufs) {
        if (b.pinned) (void)hipHostFree(b.p);
        else (void)hipFree(b.p);
    }
    ctx->pool.bufs.clear();
    for (hipStream_t st : ctx->pool.streams) (void)hipStreamDestroy(st);
    ctx->pool.streams.clear();
}

bool lockstep_eligible(mgf_scene s) { return lock_eligible(s); }

// The extrapolation layers of every scene from its offsets (in Q.offs_host):
// for each layer index, one k_multi_layer launch over the scenes whose layer
// is larger than kBatchTailMax entries, then one launch of every scene's
// remaining layers (from the first layer after which none is large) in one
// work-group each.  A scene's layers run in order: its large layers are a
// prefix, and launches on one stream run in order.
mgf_status launch_layers(hipStream_t st, const std::vector<mgf_scene>& sc, const PhasePlan& Q,
                         std::vector<unsigned char>& host, unsigned char*& dev, size_t& dev_cap) {
    constexpr uint32_t kBatchTailMax = 4096;
    const size_t n = sc.size();
    const uint32_t epb = layer_entries_per_block();
    const size_t tb = tail_arg_bytes();
    std::vector<std::vector<uint32_t>> maps;
    std::vector<uint32_t> tail_map;
    host.assign((n * tb + 255) / 256 * 256, 0);
    for (size_t i = 0; i < n; ++i) {
        const int32_t cap = Q.cap[i];
        if (cap <= 0) continue;
        const uint32_t* o = Q.offs_host + Q.offs_at[i];
        const int32_t top = int32_t(o[cap + 2]);
        int32_t first = top + 1;
        while (first > 1 && o[first] - o[first - 1] <= kBatchTailMax) --first;
        if (maps.size() < size_t(first)) maps.resize(size_t(first));
        for (int32_t L = 1; L < first; ++L) {
            const uint32_t cnt = o[L + 1] - o[L];
            for (uint32_t b = 0; b * epb < cnt; ++b) maps[size_t(L)].push_back((uint32_t(i) << kMapShift) | b);
        }
        if (first <= top) {
            mgf_scene s = sc[i];
            tail_args(host.data() + i * tb, s->g, s->u[s->cur], s->v[s->cur], s->w[s->cur], s->list, s->lmask,
                      s->counts, ext_max_distance(s->g), first, top);
            tail_map.push_back(uint32_t(i) << kMapShift);
        }
    }
    std::vector<size_t> map_at(maps.size(), 0);
    for (size_t L = 0; L < maps.size(); ++L) {
        map_at[L] = host.size();
        const unsigned char* p = reinterpret_cast<const unsigned char*>(maps[L].data());
        host.insert(host.end(), p, p + 4 * maps[L].size());
    }
    const size_t tail_map_at = host.size();
    const unsigned char* tp = reinterpret_cast<const unsigned char*>(tail_map.data());
    host.insert(host.end(), tp, tp + 4 * tail_map.size());
    if (host.size() > dev_cap) {
        MGF_HIP(hipStreamSynchronize(st));
        pool_put(sc[0]->ctx, dev);
        dev = nullptr;
        dev_cap = 0;
        const size_t want = host.size() + host.size() / 2 + 4096;
        dev = static_cast<unsigned char*>(pool_get(sc[0]->ctx, want, false));
        if (!dev) {
            return fail(MGF_ERR_NOMEM, "lockstep layers: hipMalloc failed");
        }
        dev_cap = want;
    }
    MGF_HIP(hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, st));
    for (size_t L = 1; L < maps.size(); ++L)
        MGF_HIP(launch_multi_layer(st, Q.dev + Q.layer_tab, reinterpret_cast<const uint32_t*>(dev + map_at[L]),
                                   uint32_t(maps[L].size()), int32_t(L)));
    MGF_HIP(launch_multi(st, kMkExtTail, dev, reinterpret_cast<const uint32_t*>(dev + tail_map_at),
                         uint32_t(tail_map.size()), 0, nullptr));
    return MGF_OK;
}

mgf_status batch_lockstep(mgf_context ctx, const std::vector<mgf_scene>& sc, int32_t nsteps, int32_t workers,
                          std::vector<mgf_step_stats>& stats, std::vector<mgf_status>& status,
                          std::vector<std::string>& errs, double* step_ms) {
    (void)workers;
    const int n = int(sc.size());
    stats.assign(size_t(n), mgf_step_stats{});
    status.assign(size_t(n), MGF_OK);
    errs.assign(size_t(n), std::string());
    if (n == 0) return MGF_OK;
    MGF_HIP(hipSetDevice(ctx->device));
    LockPlan P;
    P.ctx = ctx;
    P.sc = sc;
    P.stream = pool_stream(ctx);
    if (!P.stream) return fail(MGF_ERR_DEVICE, "hipStreamCreate failed");
    hipEvent_t ev = nullptr;
    std::vector<std::pair<std::vector<uint8_t>, std::unique_ptr<PhasePlan>>> phases;
    std::vector<unsigned char> step_host;   // this step's layer maps and tail table
    unsigned char* step_dev = nullptr;
    size_t step_cap = 0;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(P.stream);
        if (ev) (void)hipEventDestroy(ev);
        pool_put(ctx, step_dev);
        phases.clear();
        pool_stream_put(ctx, P.stream);
        P.stream = nullptr;
    };
    auto fin = [&](mgf_status st) {
        cleanup();
        return st;
    };
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return fin(fail(MGF_ERR_DEVICE, "lockstep batch: hipEventCreate failed"));
    }
    // with timing on (any scene), the batch step's phases are timed: before
    // the solve, the lockstep solve, after the solve
    bool timing = false;
    for (mgf_scene s : sc) timing = timing || s->timing;
    for (int i = 0; i < n; ++i) {
        stats_reset(&stats[size_t(i)]);
        mgf_scene s = sc[size_t(i)];
        mgf_status st = lock_prepare(s);
        if (st == MGF_OK) st = div_begin(s);
        // what the caller queued on the scene's stream comes first
        if (st == MGF_OK && (hipEventRecord(ev, s->stream) != hipSuccess ||
                             hipStreamWaitEvent(P.stream, ev, 0) != hipSuccess)) {
            (void)hipGetLastError();
            st = fail(MGF_ERR_DEVICE, "hipEventRecord failed");
        }
        if (st != MGF_OK) return fin(st);
    }
    mgf_status result = build_plan(P);
    if (result == MGF_OK) result = capture_iteration(P);
    if (result != MGF_OK) return fin(result);
    if (nsteps >= 2) {
        // Every scene flips its buffers each step, so the steps alternate
        // between two phase plans.  Both are built now, the second on a
        // helper thread with a stream of its own: built at the second step,
        // it would hold the device idle for the host's whole build.  The
        // scenes' hierarchies and schedules exist (lock_prepare) and nothing
        // changes them while the two builds read them.
        std::vector<uint8_t> ca(static_cast<size_t>(n)), cb(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            ca[size_t(i)] = uint8_t(sc[size_t(i)]->cur);
            cb[size_t(i)] = uint8_t(1 - sc[size_t(i)]->cur);
        }
        phases.emplace_back(ca, std::make_unique<PhasePlan>());
        PhasePlan* qa = phases.back().second.get();
        phases.emplace_back(cb, std::make_unique<PhasePlan>());
        PhasePlan* qb = phases.back().second.get();
        hipStream_t sb = pool_stream(ctx);
        if (!sb) return fin(fail(MGF_ERR_DEVICE, "hipStreamCreate failed"));
        mgf_status rb = MGF_OK;
        std::string eb;
        auto build_b = [&]() {
            if (hipSetDevice(ctx->device) != hipSuccess) {
                (void)hipGetLastError();
                rb = fail(MGF_ERR_DEVICE, "lockstep phases: hipSetDevice failed on the helper thread");
            } else {
                rb = build_phase(sc, cb, sb, *qb);
            }
            if (rb != MGF_OK) eb = mgf_last_error();
        };
        // a helper that cannot be started leaves the build to this thread:
        // nothing crosses the C boundary as an exception
        std::thread helper;
        try {
            helper = std::thread(build_b);
        } catch (...) {
        }
        result = build_phase(sc, ca, P.stream, *qa);
        if (helper.joinable()) helper.join();
        else if (result == MGF_OK) build_b();
        pool_stream_put(ctx, sb);
        if (result == MGF_OK && rb != MGF_OK) {
            set_error(eb);
            result = rb;
        }
        if (result != MGF_OK) return fin(result);
    }
    std::vector<uint8_t> cur(static_cast<size_t>(n), 0);
    for (int32_t step = 0; step < nsteps && result == MGF_OK; ++step) {
        const double t0 = now_ms();
        // boundaries that changed (added or moving solids) on the scenes'
        // streams, then everything else on the batch stream
        for (int i = 0; i < n; ++i) {
            mgf_scene s = sc[size_t(i)];
            bool moving = false;
            for (auto& so : s->solids)
                if (so.vel[0] != 0.0f || so.vel[1] != 0.0f || so.vel[2] != 0.0f) moving = true;
            if (!s->boundary_dirty && !moving) {
                result = refresh_boundaries(s);   // static solids: the translations only
                if (result != MGF_OK) return fin(result);
                continue;
            }
            // after the batch stream's work so far, on the scene's stream
            if (hipEventRecord(ev, P.stream) != hipSuccess || hipStreamWaitEvent(s->stream, ev, 0) != hipSuccess) {
                (void)hipGetLastError();
                return fin(fail(MGF_ERR_DEVICE, "hipEventRecord failed"));
            }
            result = refresh_boundaries(s);
            if (result != MGF_OK) return fin(result);
            if (hipEventRecord(ev, s->stream) != hipSuccess || hipStreamWaitEvent(P.stream, ev, 0) != hipSuccess) {
                (void)hipGetLastError();
                return fin(fail(MGF_ERR_DEVICE, "hipEventRecord failed"));
            }
        }
        for (int i = 0; i < n; ++i) cur[size_t(i)] = uint8_t(sc[size_t(i)]->cur);
        PhasePlan* Q = nullptr;
        for (auto& q : phases)
            if (q.first == cur) Q = q.second.get();
        if (!Q) {
            phases.emplace_back(cur, std::make_unique<PhasePlan>());
            Q = phases.back().second.get();
            result = build_phase(sc, cur, P.stream, *Q);
            if (result != MGF_OK) return fin(result);
        }
        result = run_seq_at(P.stream, Q->dev, Q->pre, nullptr);
        if (result != MGF_OK) return fin(result);
        for (int i = 0; i < n; ++i) {
            mgf_scene s = sc[size_t(i)];
            s->cur = 1 - s->cur;
            s->wet_flat_cells = s->wet0_cells = -1.0;
            s->x64_valid = false;
        }
        if (timing) MGF_HIP(hipStreamSynchronize(P.stream));
        const double t1 = now_ms();
        result = lock_solve(P, stats);
        if (result != MGF_OK) return fin(result);
        const double t2 = now_ms();
        result = run_seq_at(P.stream, Q->dev, Q->post, nullptr);
        if (result != MGF_OK) return fin(result);
        for (int i = 0; i < n; ++i) {
            mgf_scene s = sc[size_t(i)];
            if (s->div_check)
                MGF_HIP(launch_div_check(P.stream, s->g, s->markers, s->u[s->cur], s->v[s->cur], s->w[s->cur], s->b,
                                         s->red.partials, kDivBlocks, s->div_dev));
        }
        // the layers: each scene's offsets to the host, then per layer index
        // one launch over every scene whose layer is large, then each
        // scene's tail of small layers in one work-group (all scenes in one
        // launch)
        if (Q->offs_words > 0) {
            MGF_HIP(hipMemcpyAsync(Q->offs_host, Q->dev, sizeof(uint32_t) * Q->offs_words, hipMemcpyDeviceToHost,
                                   P.stream));
            MGF_HIP(hipStreamSynchronize(P.stream));
            result = launch_layers(P.stream, sc, *Q, step_host, step_dev, step_cap);
            if (result != MGF_OK) return fin(result);
        }
        MGF_HIP(hipStreamSynchronize(P.stream));
        for (int i = 0; i < n; ++i) {
            mgf_step_stats& a = stats[size_t(i)];
            a.extrapolation_layers = Q->cap[size_t(i)] > 0 ? int32_t(Q->offs_host[Q->offs_at[size_t(i)] + size_t(Q->cap[size_t(i)]) + 2]) : 0;
            a.steps += 1;
            if (!std::isfinite(a.residual_rel) && status[size_t(i)] == MGF_OK) {
                status[size_t(i)] = MGF_ERR_NONFINITE;
                errs[size_t(i)] = "pressure solve produced NaN/Inf";
            }
        }
        const double dt_ms = now_ms() - t0;
        if (timing)
            for (int i = 0; i < n; ++i) {
                mgf_step_stats& a = stats[size_t(i)];
                a.ms_advect += t1 - t0;
                a.ms_pressure += t2 - t1;
                a.ms_extrapolate += dt_ms - (t2 - t0);
                a.ms_total += dt_ms;
            }
        if (step_ms)
            for (int i = 0; i < n; ++i) step_ms[size_t(i) * size_t(nsteps) + size_t(step)] = dt_ms;
        for (int i = 0; i < n; ++i)
            if (status[size_t(i)] != MGF_OK) result = status[size_t(i)];
    }
    for (int i = 0; i < n; ++i) {
        mgf_scene s = sc[size_t(i)];
        mgf_status st = div_end(s, &stats[size_t(i)]);
        if (st != MGF_OK && status[size_t(i)] == MGF_OK) {
            status[size_t(i)] = st;
            errs[size_t(i)] = mgf_last_error();
        }
        if (status[size_t(i)] == MGF_OK && !stats[size_t(i)].converged) {
            status[size_t(i)] = MGF_ERR_NOT_CONVERGED;
            errs[size_t(i)] = "a pressure solve stopped before its tolerance";
        }
    }
    return fin(result);
}

mgf_status scene_prepare(mgf_scene s) {
    if (s->desc.precision == MGF_PRECISION_MIXED && s->desc.preconditioner != MGF_PRECOND_IDENTITY) {
        MGF_HIP(hipSetDevice(s->ctx->device));
        TRY(ensure_precond(s));
    }
    return MGF_OK;
}

}  // namespace mgf

// ===========================================================================
extern "C" {

const char* mgf_status_string(mgf_status s) {
    switch (s) {
        case MGF_OK: return "ok";
        case MGF_ERR_INVALID: return "invalid argument";
        case MGF_ERR_DEVICE: return "device error";
        case MGF_ERR_NOMEM: return "out of device memory";
        case MGF_ERR_NOT_CONVERGED: return "not converged";
        case MGF_ERR_NONFINITE: return "non-finite value";
        case MGF_ERR_UNSUPPORTED: return "unsupported";
        case MGF_ERR_BUDGET: return "HBM budget full";
        default: return "unknown status";
    }
}

const char* mgf_last_error(void) { return g_last_error.c_str(); }

mgf_status mgf_init(int32_t device, mgf_context* out) {
    if (!out) return fail(MGF_ERR_INVALID, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(MGF_ERR_DEVICE, "no HIP device");
    }
    if (device < 0 || device >= n) return fail(MGF_ERR_INVALID, "device index out of range");
    MGF_HIP(hipSetDevice(device));
    auto* c = new mgf_context_s;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MGF_ERR_DEVICE, "hipStreamCreate failed");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        c->lds_per_block = prop.sharedMemPerBlock;
    }
    na_setup(device);   // the no-allocation entry point's scratch on this device
    *out = c;
    return MGF_OK;
}

mgf_status mgf_batch_set_lockstep(mgf_context ctx, int32_t enabled) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    ctx->lockstep = enabled != 0;
    return MGF_OK;
}

mgf_status mgf_shutdown(mgf_context ctx) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    pool_release(ctx);
    delete ctx;
    return MGF_OK;
}

mgf_status mgf_scene_desc_defaults(mgf_scene_desc* d, int32_t nx, int32_t ny, int32_t nz,
                                   const float origin[3], const float spacing[3], float dt) {
    if (!d) return fail(MGF_ERR_INVALID, "null descriptor");
    std::memset(d, 0, sizeof(*d));
    d->nx = nx; d->ny = ny; d->nz = nz;
    for (int k = 0; k < 3; ++k) {
        d->origin[k] = origin ? origin[k] : 0.0f;
        d->spacing[k] = spacing ? spacing[k] : 1.0f;
    }
    d->dt = dt;
    d->density = 1.0f;
    d->gravity[1] = -9.8f;
    d->rk_order = MGF_RK3;
    d->tol_mode = MGF_TOL_ABS_SQRT_NORM;
    d->tolerance = 1e-3;
    d->max_iterations = 0;
    d->precision = MGF_PRECISION_PLAIN;
    d->preconditioner = MGF_PRECOND_IDENTITY;
    d->extrapolation_layers = 1000;
    d->max_refinements = 8;
    d->inner_tolerance = 1e-3;
    d->mg_sweeps = 3;
    d->mg_omega = 0.9f;
    return MGF_OK;
}

mgf_status mgf_scene_create(mgf_context ctx, const mgf_scene_desc* desc, mgf_scene* out) {
    if (!ctx || !out) return fail(MGF_ERR_INVALID, "null context or output");
    *out = nullptr;
    TRY(validate_desc(desc));
    MGF_HIP(hipSetDevice(ctx->device));
    auto* s = new mgf_scene_s;
    s->ctx = ctx;
    shape_scene(s, desc);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        s->stream = nullptr;
        mgf_scene_destroy(s);
        return fail(MGF_ERR_DEVICE, "hipStreamCreate failed");
    }
    mgf_status st = alloc_scene(s);
    if (st != MGF_OK) {
        mgf_scene_destroy(s);
        return st;
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&s->host_scratch), sizeof(double) * kMaxBlocks * 4) !=
        hipSuccess) {
        (void)hipGetLastError();
        mgf_scene_destroy(s);
        return fail(MGF_ERR_NOMEM, "hipHostMalloc failed");
    }
    hipStream_t sm = s->stream;
    // resize() zero-fills every field (MPhysics/ScalarGrid.cpp:33-37)
    hipError_t e = hipSuccess;
    auto F = [&](float* p, size_t n) {
        if (e == hipSuccess) e = launch_fill(sm, p, n, 0.0f);
    };
    for (int k = 0; k < 2; ++k) {
        F(s->u[k], s->nfx); F(s->v[k], s->nfy); F(s->w[k], s->nfz); F(s->phi[k], s->ncell);
    }
    F(s->solid_total, s->ncell);
    F(s->us, s->nfx); F(s->vs, s->nfy); F(s->ws, s->nfz);
    F(s->x, s->ncell); F(s->b, s->ncell);
    if (e == hipSuccess) e = hipMemsetAsync(s->solid_mask, 0, s->ncell, sm);
    if (e == hipSuccess) e = hipMemsetAsync(s->markers, 0, s->ncell, sm);
    if (e == hipSuccess) e = hipStreamSynchronize(sm);
    if (e != hipSuccess) {
        mgf_scene_destroy(s);
        return fail(MGF_ERR_DEVICE, std::string("scene init: ") + hipGetErrorString(e));
    }
    *out = s;
    return MGF_OK;
}

mgf_status mgf_scene_destroy(mgf_scene s) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    (void)hipSetDevice(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : s->allocations) (void)hipFree(p);
    for (auto& so : s->solids) (void)hipFree(so.sdf0);
    if (s->host_scratch) (void)hipHostFree(s->host_scratch);
    if (s->host_offs) (void)hipHostFree(s->host_offs);
    s->sink.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return MGF_OK;
}

mgf_status mgf_scene_get_desc(mgf_scene s, mgf_scene_desc* out) {
    if (!s || !out) return fail(MGF_ERR_INVALID, "null argument");
    *out = s->desc;
    return MGF_OK;
}

mgf_status mgf_scene_set_params(mgf_scene s, const mgf_scene_desc* d) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    TRY(validate_desc(d));
    if (d->nx != s->desc.nx || d->ny != s->desc.ny || d->nz != s->desc.nz)
        return fail(MGF_ERR_INVALID, "resolution cannot change");
    for (int k = 0; k < 3; ++k)
        if (d->origin[k] != s->desc.origin[k] || d->spacing[k] != s->desc.spacing[k])
            return fail(MGF_ERR_INVALID, "origin/spacing cannot change");
    s->desc = *d;
    return MGF_OK;
}

mgf_status mgf_scene_device_bytes(mgf_scene s, int64_t* out) {
    if (!s || !out) return fail(MGF_ERR_INVALID, "null argument");
    *out = s->device_bytes + int64_t(s->solids.size() * s->ncell * sizeof(float));
    return MGF_OK;
}

mgf_status mgf_scene_add_solid(mgf_scene s, const float* sdf, int64_t count, const float velocity[3]) {
    if (!s || !sdf) return fail(MGF_ERR_INVALID, "null argument");
    if (size_t(count) != s->ncell) return fail(MGF_ERR_INVALID, "solid SDF must have nx*ny*nz values");
    if (s->solids.size() >= 64) return fail(MGF_ERR_UNSUPPORTED, "at most 64 solids");
    MGF_HIP(hipSetDevice(s->ctx->device));
    Solid so;
    if (hipMalloc(reinterpret_cast<void**>(&so.sdf0), s->ncell * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MGF_ERR_NOMEM, "hipMalloc of a solid SDF failed");
    }
    if (hipMemcpy(so.sdf0, sdf, s->ncell * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(so.sdf0);
        return fail(MGF_ERR_DEVICE, "copy of a solid SDF failed");
    }
    for (int k = 0; k < 3; ++k) so.vel[k] = velocity ? velocity[k] : 0.0f;
    s->solids.push_back(so);
    s->boundary_dirty = true;
    s->solids_unbuilt = true;
    return MGF_OK;
}

mgf_status mgf_scene_fluid_box(mgf_scene s, const float mn[3], const float mx[3]) {
    if (!s || !mn || !mx) return fail(MGF_ERR_INVALID, "null argument");
    hipStream_t st = s->stream;
    MGF_HIP(launch_fluid_box(st, s->g, mn, mx, s->phi[s->cur]));
    // generateFluidDomainFromBBox copies domain 1 into domain 2
    MGF_HIP(hipMemcpyAsync(s->phi[1 - s->cur], s->phi[s->cur], s->ncell * sizeof(float),
                           hipMemcpyDeviceToDevice, st));
    return sync(s);
}

mgf_status mgf_scene_field_count(mgf_scene s, int32_t field, int64_t* out) {
    if (!s || !out || field < 0 || field >= MGF_FIELD_COUNT) return fail(MGF_ERR_INVALID, "bad argument");
    *out = int64_t(field_count(s, field));
    return MGF_OK;
}

}  // extern "C"

namespace {

// Upload from host (kind = H2D) or device (D2D) memory; the same side effects
// either way.
mgf_status upload_field(mgf_scene s, int32_t field, const float* src, int64_t count, hipMemcpyKind kind) {
    if (!s || !src || field < 0 || field >= MGF_FIELD_COUNT) return fail(MGF_ERR_INVALID, "bad argument");
    if (count < 0 || size_t(count) != field_count(s, field)) return fail(MGF_ERR_INVALID, "element count mismatch");
    MGF_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->stream;
    float* dst = nullptr;
    const int c = s->cur;
    switch (field) {
        case MGF_FIELD_VEL_X: dst = s->u[c]; break;
        case MGF_FIELD_VEL_Y: dst = s->v[c]; break;
        case MGF_FIELD_VEL_Z: dst = s->w[c]; break;
        case MGF_FIELD_FLUID_DOMAIN: dst = s->phi[c]; break;
        case MGF_FIELD_SOLID_SDF: dst = s->solid_total; break;
        case MGF_FIELD_SOLID_VEL_X: dst = s->us; break;
        case MGF_FIELD_SOLID_VEL_Y: dst = s->vs; break;
        case MGF_FIELD_SOLID_VEL_Z: dst = s->ws; break;
        case MGF_FIELD_COLOR_X: case MGF_FIELD_COLOR_Y: case MGF_FIELD_COLOR_Z:
            if (!s->color[0][0]) {
                // both buffers of the three components, zero-filled as resize() fills
                for (int b = 0; b < 2; ++b)
                    for (int k = 0; k < 3; ++k) {
                        TRY(dalloc(s, &s->color[b][k], s->ncell));
                        MGF_HIP(launch_fill(st, s->color[b][k], s->ncell, 0.0f));
                    }
            }
            dst = s->color[c][field - MGF_FIELD_COLOR_X];
            break;
        default: return fail(MGF_ERR_INVALID, "field is download-only");
    }
    MGF_HIP(hipMemcpyAsync(dst, src, size_t(count) * sizeof(float), kind, st));
    if (field == MGF_FIELD_SOLID_SDF) MGF_HIP(launch_solid_mask(st, s->ncell, s->solid_total, s->solid_mask));
    // The next boundary refresh (mgf_step) recomputes these from the solids,
    // as CBoundarys::updateBoundarys overwrites setSolidVelField.
    if (field >= MGF_FIELD_SOLID_SDF && field <= MGF_FIELD_SOLID_VEL_Z) s->boundary_dirty = true;
    return sync(s);
}

mgf_status download_field(mgf_scene s, int32_t field, float* out, int64_t count, hipMemcpyKind kind) {
    if (!s || !out || field < 0 || field >= MGF_FIELD_COUNT) return fail(MGF_ERR_INVALID, "bad argument");
    if (count < 0 || size_t(count) != field_count(s, field)) return fail(MGF_ERR_INVALID, "element count mismatch");
    MGF_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->stream;
    const int c = s->cur;
    const float* src = nullptr;
    float* tmp = nullptr;
    switch (field) {
        case MGF_FIELD_VEL_X: src = s->u[c]; break;
        case MGF_FIELD_VEL_Y: src = s->v[c]; break;
        case MGF_FIELD_VEL_Z: src = s->w[c]; break;
        case MGF_FIELD_FLUID_DOMAIN: src = s->phi[c]; break;
        case MGF_FIELD_SOLID_SDF: src = s->solid_total; break;
        case MGF_FIELD_SOLID_VEL_X: src = s->us; break;
        case MGF_FIELD_SOLID_VEL_Y: src = s->vs; break;
        case MGF_FIELD_SOLID_VEL_Z: src = s->ws; break;
        case MGF_FIELD_RHS: src = s->b; break;
        case MGF_FIELD_COLOR_X: case MGF_FIELD_COLOR_Y: case MGF_FIELD_COLOR_Z:
            if (!s->color[0][0]) {
                if (kind == hipMemcpyDeviceToHost) {
                    std::fill(out, out + count, 0.0f);
                    return MGF_OK;
                }
                MGF_HIP(launch_fill(st, out, size_t(count), 0.0f));
                return sync(s);
            }
            src = s->color[c][field - MGF_FIELD_COLOR_X];
            break;
        case MGF_FIELD_PRESSURE:
            if (s->x_is_64) {
                // the FP64 pressure rounded to FP32 (in the scratch vector ap)
                MGF_HIP(launch_f64_to_f32(st, s->x64, s->ap, s->ncell));
                src = s->ap;
            } else {
                src = s->x;
            }
            break;
        case MGF_FIELD_MARKERS:
            MGF_HIP(launch_u8_to_f32(st, s->markers, s->ap, s->ncell));
            src = s->ap;
            break;
        case MGF_FIELD_DIVERGENCE:
            MGF_HIP(launch_divergence(st, s->g, vel_of(s->u[c], s->v[c], s->w[c]), s->ap));
            src = s->ap;
            break;
        case MGF_FIELD_FLUID_DENSITY:
            MGF_HIP(launch_fluid_density(st, s->g, s->phi[c], s->solid_total, s->ap));
            src = s->ap;
            break;
        case MGF_FIELD_FLUID_SDF:
            if (s->desc.extrapolation_layers > 0)
                MGF_HIP(launch_fluid_sdf(st, s->g, s->phi[c], s->desc.extrapolation_layers, s->layer, s->ap));
            else   // no pass: the freshly resized field's zeros (ScalarGrid.cpp:33-37)
                MGF_HIP(launch_fill(st, s->ap, s->ncell, 0.0f));
            src = s->ap;
            break;
        case MGF_FIELD_MATRIX_A: {
            float sa[3], sg[3];
            scales(s, sa, sg);
            if (hipMalloc(reinterpret_cast<void**>(&tmp), 4 * s->ncell * sizeof(float)) != hipSuccess) {
                (void)hipGetLastError();
                return fail(MGF_ERR_NOMEM, "temporary matrix buffer");
            }
            MGF_HIP(launch_matrix_a(st, s->g, sa, s->markers, tmp));
            src = tmp;
            break;
        }
        default: return fail(MGF_ERR_INVALID, "unknown field");
    }
    hipError_t e = hipMemcpyAsync(out, src, size_t(count) * sizeof(float), kind, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) return fail(MGF_ERR_DEVICE, std::string("download: ") + hipGetErrorString(e));
    return MGF_OK;
}

}  // namespace

extern "C" {

mgf_status mgf_scene_upload(mgf_scene s, int32_t field, const float* host, int64_t count) {
    return upload_field(s, field, host, count, hipMemcpyHostToDevice);
}

mgf_status mgf_scene_download(mgf_scene s, int32_t field, float* host, int64_t count) {
    return download_field(s, field, host, count, hipMemcpyDeviceToHost);
}

mgf_status mgf_scene_upload_device(mgf_scene s, int32_t field, const float* dev, int64_t count) {
    return upload_field(s, field, dev, count, hipMemcpyDeviceToDevice);
}

mgf_status mgf_scene_download_device(mgf_scene s, int32_t field, float* dev, int64_t count) {
    return download_field(s, field, dev, count, hipMemcpyDeviceToDevice);
}

mgf_status mgf_scene_update_boundaries(mgf_scene s) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    ProbeScope probe(s);
    TRY(refresh_boundaries(s));
    return sync(s);
}

mgf_status mgf_scene_advect(mgf_scene s) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    ProbeScope probe(s);
    TRY(advect(s, false));
    return sync(s);
}

mgf_status mgf_scene_apply_forces(mgf_scene s) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    ProbeScope probe(s);
    TRY(apply_forces(s));
    return sync(s);
}

mgf_status mgf_scene_pressure_solve(mgf_scene s, mgf_step_stats* stats) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    ProbeScope probe(s);
    mgf_step_stats acc;
    stats_reset(&acc);
    TRY(div_begin(s));
    TRY(pressure(s, &acc));
    TRY(sync(s));
    TRY(div_end(s, &acc));
    if (stats) *stats = acc;
    if (!acc.converged) return fail(MGF_ERR_NOT_CONVERGED, "pressure solve stopped before its tolerance");
    return MGF_OK;
}

mgf_status mgf_scene_pressure_solve_particles(mgf_scene s, mgf_particles parts, mgf_step_stats* stats) {
    if (!s || !parts) return fail(MGF_ERR_INVALID, "null argument");
    if (parts->ctx->device != s->ctx->device) return fail(MGF_ERR_INVALID, "particles on another device");
    MGF_HIP(hipSetDevice(s->ctx->device));
    ProbeScope probe(s);
    mgf_step_stats acc;
    stats_reset(&acc);
    TRY(div_begin(s));
    TRY(pressure(s, &acc, parts));
    TRY(sync(s));
    TRY(div_end(s, &acc));
    if (stats) *stats = acc;
    if (!acc.converged) return fail(MGF_ERR_NOT_CONVERGED, "pressure solve stopped before its tolerance");
    return MGF_OK;
}

static bool same_grid(const mgf_scene_desc& d, const mgf_grid& g) {
    return d.nx == g.nx && d.ny == g.ny && d.nz == g.nz && d.origin[0] == g.origin[0] &&
           d.origin[1] == g.origin[1] && d.origin[2] == g.origin[2] && d.spacing[0] == g.spacing[0] &&
           d.spacing[1] == g.spacing[1] && d.spacing[2] == g.spacing[2];
}

mgf_status mgf_scene_apply_preconditioner(mgf_scene s, const float* r, float* z, double* rz) {
    if (!s || !r || !z || r == z) return fail(MGF_ERR_INVALID, "null or aliased buffers");
    const auto& d = s->desc;
    if (d.precision != MGF_PRECISION_MIXED || d.preconditioner == MGF_PRECOND_IDENTITY)
        return fail(MGF_ERR_UNSUPPORTED, "only the MIXED solver's JACOBI or MULTIGRID preconditioner");
    MGF_HIP(hipSetDevice(s->ctx->device));
    ProbeScope probe(s);
    hipStream_t st = s->stream;
    float sa[3], sg[3];
    scales(s, sa, sg);
    // the markers and codes of a pressure solve of the current state
    MGF_HIP(launch_markers(st, s->g, s->solid_mask, s->phi[s->cur], s->markers));
    MGF_HIP(launch_codes(st, s->g.nx, s->g.ny, s->g.nz, s->markers, s->codes));
    TRY(ensure_precond(s));
    TRY(mg_setup(s, sa));
    CgScalars hs{};
    MGF_HIP(hipMemcpyAsync(s->scal, &hs, sizeof(CgScalars), hipMemcpyHostToDevice, st));
    MGF_HIP(hipMemcpyAsync(s->r, r, sizeof(float) * s->ncell, hipMemcpyDeviceToDevice, st));
    const float* zv = nullptr;
    TRY(apply_precond(s, sa, &zv));
    MGF_HIP(hipMemcpyAsync(z, zv, sizeof(float) * s->ncell, hipMemcpyDeviceToDevice, st));
    MGF_HIP(hipMemcpyAsync(&hs, s->scal, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
    TRY(sync(s));
    if (rz) *rz = hs.rz;
    return MGF_OK;
}

mgf_status mgf_scene_generate_particles(mgf_scene s, int32_t per_cell, uint64_t seed, mgf_particles* out) {
    if (!s || !out) return fail(MGF_ERR_INVALID, "null argument");
    MGF_HIP(hipSetDevice(s->ctx->device));
    // solids added since the last step: their marker at the current translations
    if (s->solids_unbuilt) {
        TRY(rebuild_boundaries(s));
        TRY(sync(s));
    }
    mgf_grid g;
    g.nx = s->desc.nx; g.ny = s->desc.ny; g.nz = s->desc.nz;
    for (int k = 0; k < 3; ++k) {
        g.origin[k] = s->desc.origin[k];
        g.spacing[k] = s->desc.spacing[k];
    }
    return mgf_particles_generate(s->ctx, &g, s->phi[s->cur], s->solid_total, per_cell, seed, out);
}

mgf_status mgf_step_particles(mgf_scene s, mgf_psolver ps, int32_t nsteps, mgf_step_stats* stats) {
    if (!s || !ps) return fail(MGF_ERR_INVALID, "null argument");
    if (nsteps < 0) return fail(MGF_ERR_INVALID, "nsteps < 0");
    if (!same_grid(s->desc, ps->grid)) return fail(MGF_ERR_INVALID, "the solver's grid is not the scene's");
    if (ps->ctx->device != s->ctx->device) return fail(MGF_ERR_INVALID, "solver on another device");
    MGF_HIP(hipSetDevice(s->ctx->device));
    ProbeScope probe(s);
    mgf_step_stats acc;
    stats_reset(&acc);
    TRY(div_begin(s));
    for (int i = 0; i < nsteps; ++i) {
        const mgf_status r = step_particles_once(s, ps, &acc);
        if (r != MGF_OK) {
            (void)hipStreamSynchronize(s->stream);
            return r;
        }
    }
    TRY(sync(s));
    TRY(div_end(s, &acc));
    if (stats) *stats = acc;
    if (!acc.converged) return fail(MGF_ERR_NOT_CONVERGED, "a pressure solve stopped before its tolerance");
    return MGF_OK;
}

mgf_status mgf_scene_extrapolate_velocity(mgf_scene s, int32_t* layers_used) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    ProbeScope probe(s);
    int used = 0;
    TRY(extrapolate(s, s->u[s->cur], s->v[s->cur], s->w[s->cur], s->desc.extrapolation_layers, &used));
    if (layers_used) *layers_used = used;
    return sync(s);
}

mgf_status mgf_step(mgf_scene s, int32_t nsteps, mgf_step_stats* stats) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    if (nsteps < 0) return fail(MGF_ERR_INVALID, "nsteps < 0");
    MGF_HIP(hipSetDevice(s->ctx->device));
    ProbeScope probe(s);
    mgf_step_stats acc;
    stats_reset(&acc);
    TRY(div_begin(s));
    for (int i = 0; i < nsteps; ++i) TRY(step_once(s, &acc));
    TRY(sync(s));
    TRY(div_end(s, &acc));
    if (stats) *stats = acc;
    if (!acc.converged) return fail(MGF_ERR_NOT_CONVERGED, "a pressure solve stopped before its tolerance");
    return MGF_OK;
}

mgf_status mgf_scene_sync(mgf_scene s) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    return sync(s);
}

mgf_status mgf_scene_set_timing(mgf_scene s, int32_t enabled) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    s->timing = enabled != 0;
    return MGF_OK;
}

mgf_status mgf_scene_set_kernel_probes(mgf_scene s, int32_t enabled) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    s->probes = enabled != 0;
    return MGF_OK;
}

mgf_status mgf_scene_kernel_probes(mgf_scene s, mgf_kernel_probe* out, int32_t cap, int32_t* count,
                                   int32_t reset) {
    if (!s || !count || cap < 0 || (cap > 0 && !out)) return fail(MGF_ERR_INVALID, "null argument");
    MGF_HIP(hipSetDevice(s->ctx->device));
    TRY(sync(s));
    ProbeSink& k = s->sink;
    // per name, in order of first appearance
    std::vector<std::string> names;
    std::vector<double> longest;
    auto index_of = [&](const std::string& nm) {
        for (size_t q = 0; q < names.size(); ++q)
            if (names[q] == nm) return q;
        names.push_back(nm);
        longest.push_back(0.0);
        return names.size() - 1;
    };
    std::vector<size_t> idx(k.done_name.size());
    for (size_t r = 0; r < k.done_name.size(); ++r) {
        idx[r] = index_of(k.done_name[r]);
        longest[idx[r]] = std::max(longest[idx[r]], k.done_ms[r]);
    }
    std::vector<mgf_kernel_probe> agg(names.size());
    for (size_t q = 0; q < names.size(); ++q) {
        std::memset(&agg[q], 0, sizeof(mgf_kernel_probe));
        std::snprintf(agg[q].name, sizeof(agg[q].name), "%s", names[q].c_str());
    }
    for (size_t r = 0; r < k.done_name.size(); ++r) {
        mgf_kernel_probe& a = agg[idx[r]];
        a.launches += 1;
        a.ms += k.done_ms[r];
        if (k.done_ms[r] >= 0.25 * longest[idx[r]]) {
            a.busy_launches += 1;
            a.busy_ms += k.done_ms[r];
            a.busy_bytes += k.done_bytes[r];
        }
    }
    *count = int32_t(agg.size());
    for (int32_t q = 0; q < cap && size_t(q) < agg.size(); ++q) out[q] = agg[size_t(q)];
    if (reset) {
        k.done_name.clear();
        k.done_ms.clear();
        k.done_bytes.clear();
    }
    return MGF_OK;
}

mgf_status mgf_scene_set_divergence_check(mgf_scene s, int32_t enabled) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    s->div_check = enabled != 0;
    return MGF_OK;
}

mgf_status mgf_scene_set_block_skip(mgf_scene s, int32_t enabled) {
    if (!s) return fail(MGF_ERR_INVALID, "null scene");
    s->skip_dry = enabled != 0;
    return MGF_OK;
}

// ---- operator entry points --------------------------------------------------
mgf_status mgf_device_alloc(mgf_context ctx, int64_t bytes, void** out) {
    if (!ctx || !out || bytes < 0) return fail(MGF_ERR_INVALID, "bad argument");
    MGF_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(out, size_t(std::max<int64_t>(bytes, 1))) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MGF_ERR_NOMEM, "hipMalloc failed");
    }
    return MGF_OK;
}

mgf_status mgf_device_free(mgf_context ctx, void* p) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    MGF_HIP(hipFree(p));
    return MGF_OK;
}

mgf_status mgf_memcpy_h2d(mgf_context ctx, void* dst, const void* src, int64_t bytes) {
    if (!ctx || !dst || !src || bytes < 0) return fail(MGF_ERR_INVALID, "bad argument");
    MGF_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyHostToDevice, ctx->stream));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_memcpy_d2d(mgf_context ctx, void* dst, const void* src, int64_t bytes) {
    if (!ctx || !dst || !src || bytes < 0) return fail(MGF_ERR_INVALID, "bad argument");
    MGF_HIP(hipSetDevice(ctx->device));
    MGF_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToDevice, ctx->stream));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_memcpy_d2h(mgf_context ctx, void* dst, const void* src, int64_t bytes) {
    if (!ctx || !dst || !src || bytes < 0) return fail(MGF_ERR_INVALID, "bad argument");
    MGF_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToHost, ctx->stream));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

static mgf_status check_grid(const mgf_grid* g) {
    if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(MGF_ERR_INVALID, "bad grid");
    const double faces = double(g->nx + 1) * g->ny * g->nz + double(g->nx) * (g->ny + 1) * g->nz +
                         double(g->nx) * g->ny * (g->nz + 1);
    if (faces >= 1073741824.0) return fail(MGF_ERR_INVALID, "grid too large");
    for (int k = 0; k < 3; ++k)
        if (!(g->spacing[k] != 0.0f)) return fail(MGF_ERR_INVALID, "zero spacing");
    return MGF_OK;
}

mgf_status mgf_op_backtrace(mgf_context ctx, const mgf_grid* grid, const float* u, const float* v,
                            const float* w, float dt, int32_t rk_order, float* px, float* py,
                            float* pz, int64_t count) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    TRY(check_grid(grid));
    if (rk_order < 1 || rk_order > 3 || count < 0) return fail(MGF_ERR_INVALID, "bad order or count");
    MGF_HIP(launch_backtrace(ctx->stream, make_grid(grid), vel_of(u, v, w), dt, rk_order, px, py, pz, count));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_sample_cc(mgf_context ctx, const mgf_grid* grid, const float* f, const float* px,
                            const float* py, const float* pz, float* out, int64_t count) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    TRY(check_grid(grid));
    if (count < 0) return fail(MGF_ERR_INVALID, "bad count");
    MGF_HIP(launch_sample_cc(ctx->stream, make_grid(grid), f, px, py, pz, out, count));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_advect_fc(mgf_context ctx, const mgf_grid* grid, const float* iu, const float* iv,
                            const float* iw, const float* u, const float* v, const float* w, float dt,
                            int32_t rk_order, float* ou, float* ov, float* ow) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    TRY(check_grid(grid));
    if (rk_order < 1 || rk_order > 3) return fail(MGF_ERR_INVALID, "bad order");
    if (ou == iu || ov == iv || ow == iw || ou == u || ov == v || ow == w)
        return fail(MGF_ERR_INVALID, "output must not alias an input");
    MGF_HIP(launch_advect_fc(ctx->stream, make_grid(grid), vel_of(iu, iv, iw), vel_of(u, v, w), dt,
                             rk_order, ou, ov, ow, 0, nullptr, 0, nullptr));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_advect_cc(mgf_context ctx, const mgf_grid* grid, const float* in, const float* u,
                            const float* v, const float* w, float dt, int32_t rk_order, float* out) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    TRY(check_grid(grid));
    if (rk_order < 1 || rk_order > 3) return fail(MGF_ERR_INVALID, "bad order");
    if (out == in) return fail(MGF_ERR_INVALID, "output must not alias the input");
    MGF_HIP(launch_advect_cc(ctx->stream, make_grid(grid), in, vel_of(u, v, w), dt, rk_order, out));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_pressure_matvec(mgf_context ctx, const mgf_grid* grid, const float scale[3],
                                  const uint8_t* markers, const float* x, float* y) {
    if (!ctx || !scale) return fail(MGF_ERR_INVALID, "null argument");
    TRY(check_grid(grid));
    if (x == y) return fail(MGF_ERR_INVALID, "y must not alias x");
    MGF_HIP(launch_matvec(ctx->stream, make_grid(grid), scale, markers, x, y));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_apply_forces(mgf_context ctx, const mgf_grid* grid, float* u, float* v, float* w,
                               const float gravity[3], const float force[3], float dt) {
    if (!ctx || !u || !v || !w || !gravity || !force) return fail(MGF_ERR_INVALID, "null argument");
    TRY(check_grid(grid));
    MGF_HIP(hipSetDevice(ctx->device));
    const size_t nx = size_t(grid->nx), ny = size_t(grid->ny), nz = size_t(grid->nz);
    const size_t n[3] = {(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)};
    float* f[3] = {u, v, w};
    // CExternalForcesSolver::applyExternalForces (ExternalForcesSolver.cpp:50-59): gravity,
    // then the external forces, each as plusAlphaX with the unit field, skipped when every
    // component of the increment is below FLT_EPSILON (FaceCenteredVectorGrid.cpp:162-175)
    const float* incs[2] = {gravity, force};
    for (int q = 0; q < 2; ++q) {
        const float a[3] = {incs[q][0] * dt, incs[q][1] * dt, incs[q][2] * dt};
        if (std::fabs(a[0]) < MGF_FLT_EPSILON && std::fabs(a[1]) < MGF_FLT_EPSILON &&
            std::fabs(a[2]) < MGF_FLT_EPSILON)
            continue;
        for (int k = 0; k < 3; ++k) MGF_HIP(launch_add_uniform(ctx->stream, f[k], n[k], a[k]));
    }
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

// ---- particles ---------------------------------------------------------------
static bool pg_alg_ok(int32_t alg) { return alg >= MGF_PG_P2GSUM && alg <= MGF_PG_CUBIC; }

mgf_status mgf_op_particles_to_grid(mgf_context ctx, const mgf_grid* grid, int32_t field_kind, int64_t n,
                                    const float* pos, const float* values, int32_t span, int32_t offset, int32_t alg,
                                    float* field, float* weight, int32_t normalize) {
    if (!ctx || !field || !weight || (n > 0 && (!pos || !values))) return fail(MGF_ERR_INVALID, "null argument");
    TRY(check_grid(grid));
    if (field_kind < 0 || field_kind > 3 || n < 0 || span < 1 || offset < 0 || offset >= span)
        return fail(MGF_ERR_INVALID, "bad field kind, count, span or offset");
    if (!pg_alg_ok(alg) || alg == MGF_PG_G2PNEAREST)
        return fail(MGF_ERR_INVALID, "particle-to-grid algorithm must be P2GSUM, LINEAR, QUADRATIC or CUBIC");
    MGF_HIP(hipSetDevice(ctx->device));
    MGF_HIP(launch_p2g(ctx->stream, make_grid(grid), field_kind, n, pos, values, span, offset, alg, field, weight,
                       normalize));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_grid_to_particles(mgf_context ctx, const mgf_grid* grid, int32_t field_kind, int64_t n,
                                    const float* pos, const float* field, int32_t alg, float* values, int32_t span,
                                    int32_t offset) {
    if (!ctx || (n > 0 && (!pos || !field || !values))) return fail(MGF_ERR_INVALID, "null argument");
    TRY(check_grid(grid));
    if (field_kind < 0 || field_kind > 3 || n < 0 || span < 1 || offset < 0 || offset >= span)
        return fail(MGF_ERR_INVALID, "bad field kind, count, span or offset");
    if (!pg_alg_ok(alg) || alg == MGF_PG_P2GSUM)
        return fail(MGF_ERR_INVALID, "grid-to-particle algorithm must be G2PNEAREST, LINEAR, QUADRATIC or CUBIC");
    MGF_HIP(hipSetDevice(ctx->device));
    MGF_HIP(launch_g2p(ctx->stream, make_grid(grid), field_kind, n, pos, field, alg, values, span, offset));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_particle_markers(mgf_context ctx, const mgf_grid* grid, const float* solid, int64_t n,
                                   const float* pos, float* out) {
    if (!ctx || !out || (n > 0 && !pos) || n < 0) return fail(MGF_ERR_INVALID, "bad argument");
    TRY(check_grid(grid));
    MGF_HIP(hipSetDevice(ctx->device));
    MGF_HIP(launch_particle_markers(ctx->stream, make_grid(grid), solid, n, pos, out));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_axpy(mgf_context ctx, int64_t n, float a, const float* x, float* y) {
    if (!ctx || n < 0 || (n > 0 && (!x || !y))) return fail(MGF_ERR_INVALID, "bad argument");
    MGF_HIP(hipSetDevice(ctx->device));
    MGF_HIP(launch_axpy_n(ctx->stream, size_t(n), a, x, y));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_scale(mgf_context ctx, int64_t n, float a, float* x) {
    if (!ctx || n < 0 || (n > 0 && !x)) return fail(MGF_ERR_INVALID, "bad argument");
    MGF_HIP(hipSetDevice(ctx->device));
    MGF_HIP(launch_scale_n(ctx->stream, size_t(n), a, x));
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    return MGF_OK;
}

mgf_status mgf_op_advect_particles(mgf_context ctx, const mgf_grid* grid, const float* u, const float* v,
                                   const float* w, int64_t n, float* pos, float dt, float cfl, int32_t order,
                                   int32_t fix_to_grid) {
    if (!ctx || !u || !v || !w || n < 0 || (n > 0 && !pos)) return fail(MGF_ERR_INVALID, "bad argument");
    TRY(check_grid(grid));
    if (!std::isfinite(dt) || !std::isfinite(cfl)) return fail(MGF_ERR_INVALID, "dt and cfl must be finite");
    if (n == 0) return MGF_OK;
    MGF_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t m = 3 * size_t(n);
    // the sampled velocities, then the max word
    float* buf = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&buf), m * sizeof(float) + 64) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MGF_ERR_NOMEM, "particle advection scratch");
    }
    const mgf_status r = particles_advect_run(st, make_grid(grid), vel_of(u, v, w), n, pos, dt, cfl, order,
                                              fix_to_grid, buf, reinterpret_cast<uint32_t*>(buf + m));
    (void)hipStreamSynchronize(st);
    (void)hipFree(buf);
    return r;
}

mgf_status mgf_op_extrapolate_fc(mgf_context ctx, const mgf_grid* grid, float* u, float* v, float* w,
                                 int32_t layers, int32_t* layers_used) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    TRY(check_grid(grid));
    if (layers < 0) return fail(MGF_ERR_INVALID, "layers < 0");
    // A transient scene-shaped holder of the extrapolation scratch.
    MGF_HIP(hipSetDevice(ctx->device));
    mgf_scene_s tmp;
    tmp.ctx = ctx;
    tmp.stream = ctx->stream;
    tmp.g = make_grid(grid);
    const size_t nface = size_t(grid->nx + 1) * grid->ny * grid->nz +
                         size_t(grid->nx) * (grid->ny + 1) * grid->nz +
                         size_t(grid->nx) * grid->ny * (grid->nz + 1);
    mgf_status st = alloc_extrapolation(&tmp, nface);
    int used = 0;
    if (st == MGF_OK) st = extrapolate(&tmp, u, v, w, layers, &used);
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : tmp.allocations) (void)hipFree(p);
    if (tmp.host_offs) (void)hipHostFree(tmp.host_offs);
    if (layers_used) *layers_used = used;
    return st;
}

mgf_status mgf_op_extrapolate_fc_noalloc(mgf_context ctx, const mgf_grid* grid, float* u, float* v, float* w,
                                         int32_t layers, int32_t* layers_used) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    TRY(check_grid(grid));
    if (layers < 0) return fail(MGF_ERR_INVALID, "layers < 0");
    if (!u || !v || !w) return fail(MGF_ERR_INVALID, "null face array");
    const size_t nface = size_t(grid->nx + 1) * grid->ny * grid->nz + size_t(grid->nx) * (grid->ny + 1) * grid->nz +
                         size_t(grid->nx) * grid->ny * (grid->nz + 1);
    if (nface > kNaFaces)
        return fail(MGF_ERR_INVALID, "mgf_op_extrapolate_fc_noalloc: " + std::to_string(nface) +
                                         " faces exceed its scratch (" + std::to_string(kNaFaces) + " faces)");
    const GridDev g = make_grid(grid);
    const int32_t K = ext_max_distance(g);
    if (layers > 1 && (K > kNaMaxK || ext_counts_size(K) > kNaCountWords))
        return fail(MGF_ERR_INVALID, "mgf_op_extrapolate_fc_noalloc: nx + ny + nz + 3 = " + std::to_string(K) +
                                         " exceeds its scratch (" + std::to_string(kNaMaxK) + ")");
    std::string err;
    MGF_HIP(hipSetDevice(ctx->device));
    unsigned char* base = static_cast<unsigned char*>(na_ptr(ctx->device, &err));
    if (!base) return fail(MGF_ERR_DEVICE, "mgf_op_extrapolate_fc_noalloc: " + err);
    // the layer offsets come back into this stack buffer (no host allocation)
    uint32_t offs[kNaMaxK + 3];
    ExtBufs B{reinterpret_cast<uint16_t*>(base + kNaDistOff), reinterpret_cast<uint32_t*>(base + kNaListOff),
              base + kNaLmaskOff, base + kNaFmaskOff, reinterpret_cast<uint32_t*>(base + kNaCountOff), offs};
    int used = 0;
    const mgf_status st = extrapolate_run(ctx->stream, g, u, v, w, layers, B, &used);
    MGF_HIP(hipStreamSynchronize(ctx->stream));
    if (layers_used) *layers_used = st == MGF_OK ? used : 0;
    return st;
}

}  // extern "C"
