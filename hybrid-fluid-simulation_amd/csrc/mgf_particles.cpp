// mgf_particles.cpp — particle sets (CEulerParticles, MPhysics/EulerParticles.cpp)
// and the particle advection solvers CParticleInCell / CFluidImplicitParticle /
// CMixPICAndFLIP (MPhysics/ParticleInCell.cpp, FluidImplicitParticle.cpp,
// MixPICAndFLIP.cpp) behind the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mgf_internal.h"

using namespace mgf;

namespace {

mgf_status fail(mgf_status s, const std::string& msg) {
    set_error(msg);
    return s;
}

#define TRY(expr)                       \
    do {                                \
        mgf_status s_ = (expr);         \
        if (s_ != MGF_OK) return s_;    \
    } while (0)

template <class T>
mgf_status owned_alloc(std::vector<void*>& list, T** out, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MGF_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    list.push_back(p);
    *out = static_cast<T*>(p);
    return MGF_OK;
}

void free_all(std::vector<void*>& list) {
    for (void* p : list) (void)hipFree(p);
    list.clear();
}

mgf_status grid_ok(const mgf_grid* g) {
    if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(MGF_ERR_INVALID, "bad grid");
    const double faces = double(g->nx + 1) * g->ny * g->nz + double(g->nx) * (g->ny + 1) * g->nz +
                         double(g->nx) * g->ny * (g->nz + 1);
    if (faces >= 1073741824.0) return fail(MGF_ERR_INVALID, "grid too large");
    for (int k = 0; k < 3; ++k)
        if (!(g->spacing[k] != 0.0f)) return fail(MGF_ERR_INVALID, "zero spacing");
    return MGF_OK;
}

// Every buffer of a set of n particles; the three state arrays from host
// memory or zero-filled (CEulerParticles::resize, EulerParticles.cpp:20-67).
mgf_status particles_make(mgf_context ctx, int64_t n, const float* pos, const float* vel, const float* color,
                          mgf_particles* out) {
    auto* p = new mgf_particles_s;
    p->ctx = ctx;
    p->n = n;
    const size_t m = 3 * size_t(n);
    auto run = [&]() -> mgf_status {
        TRY(owned_alloc(p->allocations, &p->pos, m));
        TRY(owned_alloc(p->allocations, &p->vel, m));
        TRY(owned_alloc(p->allocations, &p->color, m));
        TRY(owned_alloc(p->allocations, &p->scalar, size_t(n)));
        TRY(owned_alloc(p->allocations, &p->vector, m));
        TRY(owned_alloc(p->allocations, &p->scratch, m));
        TRY(owned_alloc(p->allocations, &p->vmax, 64));
        hipStream_t st = ctx->stream;
        float* dst[3] = {p->pos, p->vel, p->color};
        const float* src[3] = {pos, vel, color};
        for (int k = 0; k < 3; ++k) {
            if (src[k] && m) MGF_HIP(hipMemcpyAsync(dst[k], src[k], m * sizeof(float), hipMemcpyHostToDevice, st));
            else MGF_HIP(launch_fill(st, dst[k], m, 0.0f));
        }
        MGF_HIP(launch_fill(st, p->scalar, size_t(n), 0.0f));
        MGF_HIP(launch_fill(st, p->vector, m, 0.0f));
        MGF_HIP(hipStreamSynchronize(st));
        return MGF_OK;
    };
    const mgf_status r = run();
    if (r != MGF_OK) {
        free_all(p->allocations);
        delete p;
        return r;
    }
    *out = p;
    return MGF_OK;
}

float* particle_array(mgf_particles p, int32_t which, int64_t* count) {
    switch (which) {
        case MGF_PARTICLE_POS: *count = 3 * p->n; return p->pos;
        case MGF_PARTICLE_VEL: *count = 3 * p->n; return p->vel;
        case MGF_PARTICLE_COLOR: *count = 3 * p->n; return p->color;
        case MGF_PARTICLE_SCALAR: *count = p->n; return p->scalar;
        case MGF_PARTICLE_VECTOR: *count = 3 * p->n; return p->vector;
        default: return nullptr;
    }
}

Vel vel3(const float* const v[3]) { return Vel{v[0], v[1], v[2]}; }

}  // namespace

namespace mgf {

// CEulerParticles::advectParticlesInVelField (EulerParticles.cpp:231-325):
// substeps of cfl * min(h) / max|v| (the last two shortened to end at dt),
// RK1/2/3 with the axpys of the positions, then the clamp to the grid box.
// The substep length needs max|v| on the host, one read-back per substep; a
// substep is one fused kernel that also samples the next substep's velocity.
mgf_status particles_advect_run(hipStream_t st, const GridDev& g, Vel V, int64_t n, float* pos, float dt, float cfl,
                                int order, int fix, float* scratch, uint32_t* vmax_bits) {
    if (n == 0) return MGF_OK;
    float* v1 = scratch;
    const float minh = std::min(std::min(g.hx, g.hy), g.hz);
    float t = 0.0f;
    bool finished = false;
    MGF_HIP(launch_sample_absmax(st, g, V, n, pos, v1, vmax_bits));
    for (int64_t sub = 0; !finished; ++sub) {
        if (sub > 10000000) return fail(MGF_ERR_NONFINITE, "particle advection did not reach dt");
        uint32_t bits = 0;
        MGF_HIP(hipMemcpyAsync(&bits, vmax_bits, sizeof bits, hipMemcpyDeviceToHost, st));
        MGF_HIP(hipStreamSynchronize(st));
        float vmax;
        std::memcpy(&vmax, &bits, sizeof vmax);
        float step = vmax != 0.0f ? cfl * minh / vmax : dt;
        if (t + step >= dt) {
            step = dt - t;
            finished = true;
        } else if (t + 2 * step >= dt) {
            step = float(0.5 * double(dt - t));
        }
        if (!finished && !(step > 0.0f && std::isfinite(step)))
            return fail(MGF_ERR_NONFINITE, "particle velocity gives no finite substep");
        const float half = float(0.5 * double(step)), three_q = float(0.75 * double(step));
        const float c1 = float(2.0 / 9.0 * double(step)), c2 = float(3.0 / 9.0 * double(step));
        const float c3 = float(4.0 / 9.0 * double(step));
        MGF_HIP(launch_rk_substep(st, g, V, n, pos, v1, order, step, half, three_q, c1, c2, c3, finished ? 0 : 1,
                                  vmax_bits));
        t += step;
    }
    if (fix) MGF_HIP(launch_clamp_particles(st, g, n, pos));
    return MGF_OK;
}

// One field splatted from the particles' values (span, off) onto grid `kind`,
// normalised: LINEAR through the bins, the wider kernels through the
// per-particle splat.
mgf_status splat(mgf_psolver s, hipStream_t st, const GridDev& g, int kind, const float* val, int span, int off,
                 float* out, float* weight) {
    mgf_particles p = s->p;
    // the velocities have a copy in bin order, made with the bins
    const bool sorted = val == p->vel;
    if (s->alg == MGF_PG_LINEAR)
        MGF_HIP(launch_p2g_binned(st, g, kind, s->bin_start, s->order, s->pos_s, sorted ? s->vel_s : val, span, off,
                                  sorted ? 1 : 0, s->acc, out, weight, 1));
    else
        MGF_HIP(launch_p2g(st, g, kind, p->n, p->pos, val, span, off, s->alg, out, weight, 1));
    return MGF_OK;
}

// The solver's particles binned at their current positions.  The bins are
// reused only when `reuse` says that nothing outside the library can have
// moved the particles since they were made (within one library step); a
// caller may write positions through the borrowed device pointer.
mgf_status ensure_bins(mgf_psolver s, hipStream_t st, const GridDev& g, bool reuse) {
    mgf_particles p = s->p;
    if (s->alg != MGF_PG_LINEAR || (reuse && s->bins_version == p->pos_version)) return MGF_OK;
    MGF_HIP(launch_bin_particles(st, g, p->n, p->pos, s->bin_counts, s->bin_tiles, s->bin_start, s->bin_cursor,
                                 s->order, s->pos_s, p->vel, s->vel_s));
    s->bins_version = p->pos_version;
    return MGF_OK;
}

// advect(CFaceCenteredVectorField) of the three solvers.
mgf_status psolver_velocity(mgf_psolver s, hipStream_t st, const float* const in[3], const float* const vel[3],
                            float dt, int order, int fix, float* const out[3]) {
    mgf_particles p = s->p;
    const GridDev g = make_grid(&s->grid);
    const int64_t n = p->n;
    // the grid change since the last output (FluidImplicitParticle.cpp:98-107)
    if (s->kind != MGF_PSOLVER_PIC) {
        for (int k = 0; k < 3; ++k) {
            const size_t nf = pg_field_count(g, k + 1);
            if (s->vel_seeded) MGF_HIP(launch_flip_delta(st, nf, in[k], s->delta[k]));
            else MGF_HIP(hipMemcpyAsync(s->delta[k], in[k], nf * sizeof(float), hipMemcpyDeviceToDevice, st));
        }
        s->vel_seeded = true;
    }
    // the particle velocities: PIC transferVelField2Particles (ParticleInCell.cpp:101);
    // FLIP + the grid change (FluidImplicitParticle.cpp:109-110); MIX blends the
    // two (MixPICAndFLIP.cpp:114-129)
    const float* const* delta = s->kind == MGF_PSOLVER_PIC ? nullptr : s->delta;
    MGF_HIP(launch_transfer_vel(st, g, n, p->pos, in, delta, s->kind, s->alg, float(1.0 - double(s->mixing)),
                                s->mixing, p->vel));
    const mgf_status adv = particles_advect_run(st, g, vel3(vel), n, p->pos, dt, s->cfl, order, fix, p->scratch,
                                                p->vmax);
    ++p->pos_version;
    TRY(adv);
    // transferParticlesVel2Field, normalised by the weights: a gather over the
    // particles binned at their new positions
    TRY(ensure_bins(s, st, g, false));
    for (int k = 0; k < 3; ++k) TRY(splat(s, st, g, k + 1, p->vel, 3, k, out[k], s->weight[k]));
    if (s->kind != MGF_PSOLVER_PIC)   // m_DeltaGridVelField = voOutputField
        for (int k = 0; k < 3; ++k)
            MGF_HIP(hipMemcpyAsync(s->delta[k], out[k], pg_field_count(g, k + 1) * sizeof(float),
                                   hipMemcpyDeviceToDevice, st));
    return MGF_OK;
}

// advect(CCellCenteredScalarField) (ParticleInCell.cpp:58-75).
mgf_status psolver_scalar(mgf_psolver s, hipStream_t st, const float* in, float* out, bool reuse_bins) {
    mgf_particles p = s->p;
    const GridDev g = make_grid(&s->grid);
    if (!s->scalar_seeded) MGF_HIP(launch_g2p(st, g, 0, p->n, p->pos, in, s->alg, p->scalar, 1, 0));
    s->scalar_seeded = true;
    TRY(ensure_bins(s, st, g, reuse_bins));
    TRY(splat(s, st, g, 0, p->scalar, 1, 0, out, s->weight[0]));
    return MGF_OK;
}

// advect(CCellCenteredVectorField) (ParticleInCell.cpp:77-94).
mgf_status psolver_vector(mgf_psolver s, hipStream_t st, const float* const in[3], float* const out[3],
                          bool reuse_bins) {
    mgf_particles p = s->p;
    const GridDev g = make_grid(&s->grid);
    if (!s->vector_seeded)
        for (int k = 0; k < 3; ++k) MGF_HIP(launch_g2p(st, g, 0, p->n, p->pos, in[k], s->alg, p->vector, 3, k));
    s->vector_seeded = true;
    TRY(ensure_bins(s, st, g, reuse_bins));
    for (int k = 0; k < 3; ++k) TRY(splat(s, st, g, 0, p->vector, 3, k, out[k], s->weight[k]));
    return MGF_OK;
}

}  // namespace mgf

extern "C" {

mgf_status mgf_particles_create(mgf_context ctx, int64_t n, const float* pos, const float* vel, const float* color,
                                mgf_particles* out) {
    if (!ctx || !out || n < 0) return fail(MGF_ERR_INVALID, "bad argument");
    if (n > (int64_t(1) << 31)) return fail(MGF_ERR_INVALID, "more than 2^31 particles");
    MGF_HIP(hipSetDevice(ctx->device));
    return particles_make(ctx, n, pos, vel, color, out);
}

mgf_status mgf_particles_generate(mgf_context ctx, const mgf_grid* grid, const float* fluid_sdf,
                                  const float* solid_sdf, int32_t per_cell, uint64_t seed, mgf_particles* out) {
    if (!ctx || !out || !fluid_sdf || per_cell < 0) return fail(MGF_ERR_INVALID, "bad argument");
    TRY(grid_ok(grid));
    MGF_HIP(hipSetDevice(ctx->device));
    const GridDev g = make_grid(grid);
    const size_t ncell = size_t(grid->nx) * grid->ny * grid->nz;
    const uint32_t nt = gen_tiles(ncell);
    std::vector<void*> tmp;
    uint32_t* tiles = nullptr;
    TRY(owned_alloc(tmp, &tiles, size_t(nt) + 1));
    auto run = [&]() -> mgf_status {
        hipStream_t st = ctx->stream;
        MGF_HIP(launch_gen_count(st, ncell, fluid_sdf, solid_sdf, tiles));
        uint32_t cells = 0;
        MGF_HIP(hipMemcpyAsync(&cells, tiles + nt, sizeof cells, hipMemcpyDeviceToHost, st));
        MGF_HIP(hipStreamSynchronize(st));
        const int64_t n = int64_t(cells) * per_cell;
        if (n > (int64_t(1) << 31)) return fail(MGF_ERR_INVALID, "more than 2^31 particles");
        TRY(particles_make(ctx, n, nullptr, nullptr, nullptr, out));
        if (n > 0) {
            MGF_HIP(launch_gen_emit(st, g, fluid_sdf, solid_sdf, tiles, per_cell, seed, (*out)->pos));
            MGF_HIP(hipStreamSynchronize(st));
        }
        return MGF_OK;
    };
    *out = nullptr;
    const mgf_status r = run();
    (void)hipStreamSynchronize(ctx->stream);
    free_all(tmp);
    if (r != MGF_OK && *out) {   // no half-made set reaches the caller
        free_all((*out)->allocations);
        delete *out;
        *out = nullptr;
    }
    return r;
}

mgf_status mgf_particles_destroy(mgf_particles p) {
    if (!p) return fail(MGF_ERR_INVALID, "null particles");
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    free_all(p->allocations);
    delete p;
    return MGF_OK;
}

mgf_status mgf_particles_count(mgf_particles p, int64_t* n) {
    if (!p || !n) return fail(MGF_ERR_INVALID, "null argument");
    *n = p->n;
    return MGF_OK;
}

mgf_status mgf_particles_data(mgf_particles p, int32_t which, float** dev) {
    if (!p || !dev) return fail(MGF_ERR_INVALID, "null argument");
    int64_t count = 0;
    float* a = particle_array(p, which, &count);
    if (!a) return fail(MGF_ERR_INVALID, "unknown particle array");
    *dev = a;
    return MGF_OK;
}

mgf_status mgf_particles_upload(mgf_particles p, int32_t which, const float* host, int64_t count) {
    if (!p || (count > 0 && !host)) return fail(MGF_ERR_INVALID, "null argument");
    int64_t want = 0;
    float* a = particle_array(p, which, &want);
    if (!a) return fail(MGF_ERR_INVALID, "unknown particle array");
    if (count != want) return fail(MGF_ERR_INVALID, "count does not match the particle array");
    MGF_HIP(hipSetDevice(p->ctx->device));
    if (which == MGF_PARTICLE_POS) ++p->pos_version;
    if (count) MGF_HIP(hipMemcpyAsync(a, host, size_t(count) * sizeof(float), hipMemcpyHostToDevice, p->ctx->stream));
    MGF_HIP(hipStreamSynchronize(p->ctx->stream));
    return MGF_OK;
}

mgf_status mgf_particles_download(mgf_particles p, int32_t which, float* host, int64_t count) {
    if (!p || (count > 0 && !host)) return fail(MGF_ERR_INVALID, "null argument");
    int64_t want = 0;
    float* a = particle_array(p, which, &want);
    if (!a) return fail(MGF_ERR_INVALID, "unknown particle array");
    if (count != want) return fail(MGF_ERR_INVALID, "count does not match the particle array");
    MGF_HIP(hipSetDevice(p->ctx->device));
    if (count) MGF_HIP(hipMemcpyAsync(host, a, size_t(count) * sizeof(float), hipMemcpyDeviceToHost, p->ctx->stream));
    MGF_HIP(hipStreamSynchronize(p->ctx->stream));
    return MGF_OK;
}

mgf_status mgf_particles_advect(mgf_particles p, const mgf_grid* grid, const float* u, const float* v,
                                const float* w, float dt, float cfl, int32_t order, int32_t fix_to_grid) {
    if (!p || !u || !v || !w) return fail(MGF_ERR_INVALID, "null argument");
    TRY(grid_ok(grid));
    if (!std::isfinite(dt) || !std::isfinite(cfl)) return fail(MGF_ERR_INVALID, "dt and cfl must be finite");
    MGF_HIP(hipSetDevice(p->ctx->device));
    hipStream_t st = p->ctx->stream;
    ++p->pos_version;
    const mgf_status r = particles_advect_run(st, make_grid(grid), Vel{u, v, w}, p->n, p->pos, dt, cfl, order,
                                              fix_to_grid, p->scratch, p->vmax);
    (void)hipStreamSynchronize(st);
    return r;
}

mgf_status mgf_psolver_create(mgf_context ctx, const mgf_grid* grid, int32_t kind, mgf_particles particles,
                              float cfl, int32_t alg, float mixing, mgf_psolver* out) {
    if (!ctx || !particles || !out) return fail(MGF_ERR_INVALID, "null argument");
    TRY(grid_ok(grid));
    if (kind < MGF_PSOLVER_PIC || kind > MGF_PSOLVER_MIX) return fail(MGF_ERR_INVALID, "unknown solver kind");
    if (alg < MGF_PG_LINEAR || alg > MGF_PG_CUBIC)
        return fail(MGF_ERR_INVALID, "transfer kernel must be LINEAR, QUADRATIC or CUBIC");
    if (!std::isfinite(cfl) || !std::isfinite(mixing)) return fail(MGF_ERR_INVALID, "cfl and mixing must be finite");
    if (particles->ctx != ctx) return fail(MGF_ERR_INVALID, "particles of another context");
    MGF_HIP(hipSetDevice(ctx->device));
    auto* s = new mgf_psolver_s;
    s->ctx = ctx;
    s->p = particles;
    s->grid = *grid;
    s->kind = kind;
    s->alg = alg;
    s->cfl = cfl;
    s->mixing = mixing;
    const GridDev g = make_grid(grid);
    auto run = [&]() -> mgf_status {
        for (int k = 0; k < 3; ++k) {
            TRY(owned_alloc(s->allocations, &s->weight[k], pg_field_count(g, k + 1)));
            if (kind != MGF_PSOLVER_PIC) TRY(owned_alloc(s->allocations, &s->delta[k], pg_field_count(g, k + 1)));
        }
        if (alg == MGF_PG_LINEAR) {   // the binned two-pass splat
            const size_t nb = bin_count(g);
            TRY(owned_alloc(s->allocations, &s->bin_counts, nb));
            TRY(owned_alloc(s->allocations, &s->bin_cursor, nb));
            TRY(owned_alloc(s->allocations, &s->bin_start, nb + 1));
            TRY(owned_alloc(s->allocations, &s->bin_tiles, size_t(bin_tiles(g)) + 1));
            TRY(owned_alloc(s->allocations, &s->order, size_t(particles->n)));
            TRY(owned_alloc(s->allocations, &s->acc, p2g_acc_floats(g)));
            TRY(owned_alloc(s->allocations, &s->pos_s, 3 * size_t(particles->n)));
            TRY(owned_alloc(s->allocations, &s->vel_s, 3 * size_t(particles->n)));
        }
        MGF_HIP(hipStreamSynchronize(ctx->stream));
        return MGF_OK;
    };
    const mgf_status r = run();
    if (r != MGF_OK) {
        free_all(s->allocations);
        delete s;
        return r;
    }
    *out = s;
    return MGF_OK;
}

mgf_status mgf_psolver_destroy(mgf_psolver s) {
    if (!s) return fail(MGF_ERR_INVALID, "null solver");
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    free_all(s->allocations);
    delete s;
    return MGF_OK;
}

mgf_status mgf_psolver_advect_velocity(mgf_psolver s, const float* in_u, const float* in_v, const float* in_w,
                                       const float* vel_u, const float* vel_v, const float* vel_w, float dt,
                                       int32_t order, int32_t fix_to_grid, float* out_u, float* out_v, float* out_w) {
    if (!s || !in_u || !in_v || !in_w || !vel_u || !vel_v || !vel_w || !out_u || !out_v || !out_w)
        return fail(MGF_ERR_INVALID, "null argument");
    if (!std::isfinite(dt)) return fail(MGF_ERR_INVALID, "dt must be finite");
    const float* in[3] = {in_u, in_v, in_w};
    const float* vel[3] = {vel_u, vel_v, vel_w};
    float* out[3] = {out_u, out_v, out_w};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (out[a] == in[b] || out[a] == vel[b]) return fail(MGF_ERR_INVALID, "output aliases an input");
    MGF_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const mgf_status r = psolver_velocity(s, st, in, vel, dt, order, fix_to_grid, out);
    (void)hipStreamSynchronize(st);
    return r;
}

mgf_status mgf_psolver_advect_scalar(mgf_psolver s, const float* in, float* out) {
    if (!s || !in || !out) return fail(MGF_ERR_INVALID, "null argument");
    MGF_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const mgf_status r = psolver_scalar(s, st, in, out, false);
    (void)hipStreamSynchronize(st);
    return r;
}

mgf_status mgf_psolver_advect_vector(mgf_psolver s, const float* in_x, const float* in_y, const float* in_z,
                                     float* out_x, float* out_y, float* out_z) {
    if (!s || !in_x || !in_y || !in_z || !out_x || !out_y || !out_z) return fail(MGF_ERR_INVALID, "null argument");
    const float* in[3] = {in_x, in_y, in_z};
    float* out[3] = {out_x, out_y, out_z};
    MGF_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const mgf_status r = psolver_vector(s, st, in, out, false);
    (void)hipStreamSynchronize(st);
    return r;
}

}  // extern "C"
