// mac_oracle_capi.cpp — extern "C" surface of the CPU oracle for the tests,
// the smoke check and the benchmark's CPU baseline leg.
//
// TEST INFRASTRUCTURE ONLY: never linked into the product library.
//
// Two kinds of entry points:
//   orc_<op>_f32 / orc_<op>_f64 : one reference operator on host arrays;
//   orc_scene_*                 : a CGridFluidSolver restatement
//                                 (MPhysics/GridFluidSolver.cpp:336-537) in
//                                 FP32 (plain transliteration) or FP64.
#include "mac_oracle.hpp"

#include <chrono>
#include <memory>

using namespace mac_oracle;

extern "C" {

typedef struct orc_grid {
    int32_t nx, ny, nz;
    double origin[3];
    double spacing[3];
} orc_grid;

typedef struct orc_params {
    double dt;
    double density;
    double gravity[3];
    double external_force[3];
    int32_t rk_order;
    int32_t tol_mode;
    double tolerance;
    int64_t max_iterations;
    int32_t jacobi;
    int32_t extrapolation_layers;
} orc_params;

typedef struct orc_stats {
    int64_t cg_iterations;
    int32_t converged;
    double residual_rel;        // FP64 ||b - A x|| / ||b|| of the last solve
    double residual_recurrence; // last recurrence ||r||
    double b_norm;
    int32_t extrapolation_layers;
    int64_t fluid_cells;
    double ms_advect, ms_pressure, ms_extrapolate, ms_total;
} orc_stats;

}  // extern "C"

namespace {

template <class R> Grid<R> to_grid(const orc_grid* g) {
    Grid<R> o;
    o.nx = g->nx; o.ny = g->ny; o.nz = g->nz;
    o.ox = R(float(g->origin[0])); o.oy = R(float(g->origin[1])); o.oz = R(float(g->origin[2]));
    o.hx = R(float(g->spacing[0])); o.hy = R(float(g->spacing[1])); o.hz = R(float(g->spacing[2]));
    return o;
}

template <class R> std::vector<R> vec(const float* p, size_t n) {
    std::vector<R> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = R(p[i]);
    return v;
}
template <class R> void put(const std::vector<R>& v, float* p) {
    for (size_t i = 0; i < v.size(); ++i) p[i] = float(v[i]);
}
template <class R> void put_d(const std::vector<R>& v, double* p) {
    for (size_t i = 0; i < v.size(); ++i) p[i] = double(v[i]);
}

inline size_t nfx(const orc_grid* g) { return size_t(g->nx + 1) * g->ny * g->nz; }
inline size_t nfy(const orc_grid* g) { return size_t(g->nx) * (g->ny + 1) * g->nz; }
inline size_t nfz(const orc_grid* g) { return size_t(g->nx) * g->ny * (g->nz + 1); }
inline size_t ncc(const orc_grid* g) { return size_t(g->nx) * g->ny * g->nz; }

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// CGridFluidSolver restatement.
template <class R> struct Scene {
    Grid<R> g;
    std::vector<R> u, v, w, phi;
    std::vector<std::vector<R>> sdf0;   // original solid SDFs
    std::vector<R> svel, strans;        // 3 per solid
    std::vector<R> solid_marker, us, vs, ws;
    std::vector<std::vector<R>> sdf_cur;  // last re-sampled SDFs
    std::vector<R> last_markers, last_x, last_b, last_A;
    std::vector<R> color[3];            // CCV colour field, empty until uploaded
    bool boundary_from_solids = true;

    explicit Scene(const Grid<R>& gr) : g(gr) {
        u.assign(size_t(g.nx + 1) * g.ny * g.nz, 0);
        v.assign(size_t(g.nx) * (g.ny + 1) * g.nz, 0);
        w.assign(size_t(g.nx) * g.ny * (g.nz + 1), 0);
        phi.assign(g.cells(), 0);
        solid_marker.assign(g.cells(), 0);
        us.assign(u.size(), 0); vs.assign(v.size(), 0); ws.assign(w.size(), 0);
    }

    // CBoundarys::updateBoundarys (MPhysics/Boundary.cpp:115-150).
    void update_boundaries(R dt) {
        if (!boundary_from_solids) return;
        const size_t n = g.cells();
        // __moveBoundarys: translation += -dt * velocity (axpyReal, no skip).
        for (size_t i = 0; i < strans.size(); ++i) strans[i] = fma_r(-dt, svel[i], strans[i]);
        std::vector<std::vector<R>> cur(sdf0.size(), std::vector<R>(n));
        std::vector<const R*> ptrs;
        for (size_t s = 0; s < sdf0.size(); ++s) {
            resample_solid(g, sdf0[s].data(), &strans[3 * s], cur[s].data());
            ptrs.push_back(cur[s].data());
        }
        build_solids_marker(n, ptrs, solid_marker.data());
        build_solids_vel(g, solid_marker.data(), svel.data(), int(sdf0.size()),
                         us.data(), vs.data(), ws.data());
        sdf_cur.swap(cur);
    }

    void advect(const orc_params& p) {
        std::vector<R> nu(u.size()), nv(v.size()), nw(w.size()), nphi(phi.size());
        advect_fc(g, g, u.data(), v.data(), w.data(), u.data(), v.data(), w.data(),
                  R(p.dt), p.rk_order, nu.data(), nv.data(), nw.data());
        // CSemiLagrangian::advect(CCV) (MPhysics/SemiLagrangian.cpp:152-167): one
        // back-trace from the cell centres, then the CCS trilinear sample of each
        // component (sampleCellCenteredVectorFieldInvoker, FieldMathTool.cu:575-640);
        // the colour is advected between the velocity and the fluid domain
        // (GridFluidSolver.cpp:405-406, 413-414), all with the old velocity
        if (!color[0].empty()) {
            for (int k = 0; k < 3; ++k) {
                std::vector<R> nc(color[k].size());
                advect_cc(g, g, color[k].data(), u.data(), v.data(), w.data(), R(p.dt), p.rk_order,
                          nc.data());
                color[k].swap(nc);
            }
        }
        advect_cc(g, g, phi.data(), u.data(), v.data(), w.data(), R(p.dt), p.rk_order,
                  nphi.data());
        u.swap(nu); v.swap(nv); w.swap(nw); phi.swap(nphi);
    }

    void forces(const orc_params& p) {
        R gr[3] = {R(p.gravity[0]), R(p.gravity[1]), R(p.gravity[2])};
        R fo[3] = {R(p.external_force[0]), R(p.external_force[1]), R(p.external_force[2])};
        apply_external_forces(g, u.data(), v.data(), w.data(), gr, fo, R(p.dt));
    }

    void pressure(const orc_params& p, orc_stats* st, const R* markers_in = nullptr) {
        PressureOut<R> o = pressure_solve(g, R(p.dt), R(p.density), u.data(), v.data(),
                                          w.data(), us.data(), vs.data(), ws.data(),
                                          solid_marker.data(), phi.data(), p.tol_mode,
                                          p.tolerance, p.max_iterations, p.jacobi != 0, markers_in);
        if (st) {
            st->cg_iterations += o.cg.iterations;
            st->converged = st->converged && o.cg.converged;
            st->residual_recurrence = o.cg.r_norm;
            st->b_norm = o.cg.b_norm;
            st->residual_rel = true_residual_rel(g, o.A.data(), o.b.data(), o.x.data());
            long long fl = 0;
            for (R m : o.markers) fl += (m == 1);
            st->fluid_cells = fl;
        }
        last_markers.swap(o.markers);
        last_x.swap(o.x);
        last_b.swap(o.b);
        last_A.swap(o.A);
    }

    // _computeAdvection through a particle solver (GridFluidSolver.cpp:399-416):
    // velocity, colour, fluid domain, all with the old velocity; the boundary
    // SDF passed there has the grid's resolution, so the particles are clamped.
    void advect_particles_solver(ParticleSolver<R>& ps, const orc_params& p) {
        std::vector<R> nu(u.size()), nv(v.size()), nw(w.size()), nphi(phi.size());
        const R* in[3] = {u.data(), v.data(), w.data()};
        R* out[3] = {nu.data(), nv.data(), nw.data()};
        ps.advect_velocity(in, in, R(p.dt), p.rk_order, true, out);
        if (!color[0].empty()) {
            std::vector<R> nc[3];
            for (auto& c : nc) c.resize(g.cells());
            const R* ic[3] = {color[0].data(), color[1].data(), color[2].data()};
            R* oc[3] = {nc[0].data(), nc[1].data(), nc[2].data()};
            ps.advect_vector(ic, oc);
            for (int k = 0; k < 3; ++k) color[k].swap(nc[k]);
        }
        ps.advect_scalar(phi.data(), nphi.data());
        u.swap(nu); v.swap(nv); w.swap(nw); phi.swap(nphi);
    }

    // CPressureSolver::solvePressure with the particles' positions
    // (PressureSolver.cpp:109-134, buildFluidMarkersInvoker EulerParticlesTool.cu:366-411).
    void pressure_particles(const ParticleSolver<R>& ps, const orc_params& p, orc_stats* st) {
        std::vector<R> m(g.cells());
        particle_markers(g, solid_marker.data(), ps.np(), ps.pos.data(), m.data());
        pressure(p, st, m.data());
    }

    void step_particles(ParticleSolver<R>& ps, const orc_params& p, orc_stats* st) {
        update_boundaries(R(p.dt));
        advect_particles_solver(ps, p);
        forces(p);
        pressure_particles(ps, p, st);
        const int used = extrapolate(p);
        if (st) st->extrapolation_layers = used;
    }

    int extrapolate(const orc_params& p) {
        return extrapolate_fc(g, u.data(), v.data(), w.data(), p.extrapolation_layers);
    }

    // _onAdvanceTimeStep (MPhysics/GridFluidSolver.cpp:359-370).
    void step(const orc_params& p, orc_stats* st) {
        double t0 = now_ms();
        update_boundaries(R(p.dt));
        advect(p);
        forces(p);
        double t1 = now_ms();
        pressure(p, st);
        double t2 = now_ms();
        int used = extrapolate(p);
        double t3 = now_ms();
        if (st) {
            st->extrapolation_layers = used;
            st->ms_advect += t1 - t0;
            st->ms_pressure += t2 - t1;
            st->ms_extrapolate += t3 - t2;
            st->ms_total += t3 - t0;
        }
    }
};

struct SceneHandle {
    int f64 = 0;
    std::unique_ptr<Scene<float>> s32;
    std::unique_ptr<Scene<double>> s64;
};

template <class F> auto with_scene(SceneHandle* h, F&& f) {
    if (h->f64) return f(*h->s64);
    return f(*h->s32);
}

template <class R> std::vector<R>* field_ptr(Scene<R>& s, int field) {
    switch (field) {
        case 0: return &s.u;
        case 1: return &s.v;
        case 2: return &s.w;
        case 3: return &s.phi;
        case 4: return &s.solid_marker;
        case 5: return &s.us;
        case 6: return &s.vs;
        case 7: return &s.ws;
        case 8: return &s.last_markers;
        case 9: return &s.last_x;
        case 10: return &s.last_b;
        case 11: return &s.last_A;
        case 13: case 14: case 15:
            // the colour field exists from its first upload on (zero-filled, as
            // resize() fills, MPhysics/ScalarGrid.cpp:33-37)
            if (s.color[0].empty())
                for (auto& c : s.color) c.assign(s.g.cells(), 0);
            return &s.color[field - 13];
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

// ---- operators, FP32 ------------------------------------------------------
int orc_sample_cc_f32(const orc_grid* gr, const float* f, const float* px, const float* py,
                      const float* pz, float* out, int64_t n) {
    Grid<float> g = to_grid<float>(gr);
    for (int64_t i = 0; i < n; ++i) out[i] = sample_cc(g, f, px[i], py[i], pz[i]);
    return 0;
}

int orc_sample_fc_f32(const orc_grid* gr, const float* u, const float* v, const float* w,
                      const float* px, const float* py, const float* pz, float* ou,
                      float* ov, float* ow, int64_t n) {
    Grid<float> g = to_grid<float>(gr);
    for (int64_t i = 0; i < n; ++i) sample_fc(g, u, v, w, px[i], py[i], pz[i], ou[i], ov[i], ow[i]);
    return 0;
}

int orc_backtrace_f32(const orc_grid* gr, const float* u, const float* v, const float* w,
                      float dt, int32_t order, float* px, float* py, float* pz, int64_t n) {
    Grid<float> g = to_grid<float>(gr);
    for (int64_t i = 0; i < n; ++i) backtrace_point(g, u, v, w, dt, order, px[i], py[i], pz[i]);
    return 0;
}

int orc_advect_cc_f32(const orc_grid* gr, const float* in, const float* u, const float* v,
                      const float* w, float dt, int32_t order, float* out) {
    Grid<float> g = to_grid<float>(gr);
    advect_cc(g, g, in, u, v, w, dt, order, out);
    return 0;
}

int orc_advect_fc_f32(const orc_grid* gr, const float* iu, const float* iv, const float* iw,
                      const float* u, const float* v, const float* w, float dt,
                      int32_t order, float* ou, float* ov, float* ow) {
    Grid<float> g = to_grid<float>(gr);
    advect_fc(g, g, iu, iv, iw, u, v, w, dt, order, ou, ov, ow);
    return 0;
}

int orc_external_forces_f32(const orc_grid* gr, float* u, float* v, float* w,
                            const float grav[3], const float force[3], float dt) {
    Grid<float> g = to_grid<float>(gr);
    apply_external_forces(g, u, v, w, grav, force, dt);
    return 0;
}

int orc_divergence_f32(const orc_grid* gr, const float* u, const float* v, const float* w,
                       float* div) {
    Grid<float> g = to_grid<float>(gr);
    divergence_fc(g, u, v, w, div);
    return 0;
}

int orc_matvec_f32(const orc_grid* gr, const float* A, const float* x, float* y) {
    Grid<float> g = to_grid<float>(gr);
    matvec(g, A, x, y);
    return 0;
}

int orc_build_matrix_f32(const orc_grid* gr, const float* markers, const float scale[3],
                         float* A) {
    Grid<float> g = to_grid<float>(gr);
    build_matrix(g, markers, scale, A);
    return 0;
}

// Generic CG on a caller matrix given as dense row-major n x n (small tests
// of the solver itself, UnitTest-CuMatrixFreePCG) with an optional diagonal
// inverse preconditioner minv (nullptr = identity).
int orc_cg_dense_f32(const float* Adense, const float* minv, const float* b, float* x,
                     int64_t n, int32_t tol_mode, double tol, int64_t max_it,
                     int64_t* iterations) {
    std::vector<float> r(n), z(n), p(n), ap(n);
    auto mv = [&](const float* in, float* out) {
        for (int64_t i = 0; i < n; ++i) {
            double acc = 0;   // cuSPARSE SpMV rounding is closed: FP64 accumulate
            for (int64_t j = 0; j < n; ++j) acc += double(Adense[i * n + j]) * in[j];
            out[i] = float(acc);
        }
    };
    auto pc = [&]() {
        for (int64_t i = 0; i < n; ++i) z[i] = minv ? minv[i] * r[i] : r[i];
    };
    std::copy(b, b + n, r.begin());
    float bn = nrm2(r.data(), size_t(n));
    *iterations = 0;
    if (tol_mode == 0 ? std::sqrt(bn) < float(tol) : bn == 0) {
        std::fill(x, x + n, 0.f);
        return 0;
    }
    mv(x, r.data());
    for (int64_t i = 0; i < n; ++i) r[i] = -1.0f * r[i];
    for (int64_t i = 0; i < n; ++i) r[i] = std::fma(1.0f, b[i], r[i]);
    pc();
    std::copy(z.begin(), z.end(), p.begin());
    for (int64_t it = 0; it < max_it; ++it) {
        float rz = dot(r.data(), z.data(), size_t(n));
        mv(p.data(), ap.data());
        ++*iterations;
        float pap = dot(p.data(), ap.data(), size_t(n));
        float alpha = rz / pap;
        if (!axpy_skipped(alpha)) {
            for (int64_t i = 0; i < n; ++i) x[i] = std::fma(alpha, p[i], x[i]);
            for (int64_t i = 0; i < n; ++i) r[i] = std::fma(-alpha, ap[i], r[i]);
        }
        float rn = nrm2(r.data(), size_t(n));
        if (tol_mode == 0 ? std::sqrt(rn) < float(tol) : double(rn) <= tol * double(bn)) break;
        pc();
        float rz2 = dot(r.data(), z.data(), size_t(n));
        float beta = rz2 / rz;
        for (int64_t i = 0; i < n; ++i) p[i] = beta * p[i];
        for (int64_t i = 0; i < n; ++i) p[i] = std::fma(1.0f, z[i], p[i]);
    }
    return 0;
}

int orc_extrapolate_grid_f32(float* data, int32_t nx, int32_t ny, int32_t nz, int32_t layers,
                             int32_t* layer_ids) {
    std::vector<int32_t> L;
    int used = extrapolate_grid(data, nx, ny, nz, layers, &L);
    if (layer_ids) std::copy(L.begin(), L.end(), layer_ids);
    return used;
}

int orc_extrapolate_fc_f32(const orc_grid* gr, float* u, float* v, float* w, int32_t layers) {
    Grid<float> g = to_grid<float>(gr);
    return extrapolate_fc(g, u, v, w, layers);
}

// ---- particles, FP32 ------------------------------------------------------
// field_kind: 0 = cell-centred grid `gr`; 1, 2, 3 = the x, y, z face grid of `gr`.
static PGrid<float> pgrid_of(const orc_grid* gr, int field_kind) {
    Grid<float> g = to_grid<float>(gr);
    return field_kind == 0 ? cell_grid(g) : face_grid(g, field_kind - 1);
}

int orc_p2g_f32(const orc_grid* gr, int32_t field_kind, int64_t np, const float* pos, const float* val,
                int32_t span, int32_t off, int32_t alg, float* field, float* weight, int32_t normalize) {
    p2g(pgrid_of(gr, field_kind), size_t(np), pos, val, span, off, alg, field, weight, normalize != 0);
    return 0;
}

int orc_g2p_f32(const orc_grid* gr, int32_t field_kind, int64_t np, const float* pos, const float* field,
                int32_t alg, float* val, int32_t span, int32_t off) {
    g2p(pgrid_of(gr, field_kind), size_t(np), pos, field, alg, val, span, off);
    return 0;
}

int orc_particle_markers_f32(const orc_grid* gr, const float* solid, int64_t np, const float* pos, float* out,
                             int32_t counts_only) {
    Grid<float> g = to_grid<float>(gr);
    if (counts_only) particle_counts(g, size_t(np), pos, out);
    else particle_markers(g, solid, size_t(np), pos, out);
    return 0;
}

int orc_advect_particles_f32(const orc_grid* gr, const float* u, const float* v, const float* w, int64_t np,
                             float* pos, float dt, float cfl, int32_t order, int32_t fix) {
    advect_particles(to_grid<float>(gr), u, v, w, size_t(np), pos, dt, cfl, order, fix != 0);
    return 0;
}

int orc_fluid_density_f32(const orc_grid* gr, const float* phi, const float* solid, float* out) {
    Grid<float> g = to_grid<float>(gr);
    build_fluid_density(g, phi, solid, out);
    return 0;
}

int orc_fluid_sdf_f32(const orc_grid* gr, const float* phi, float* out, int32_t passes) {
    Grid<float> g = to_grid<float>(gr);
    build_fluid_sdf(g, phi, out, passes);
    return 0;
}

// CPressureSolver::solvePressure on host arrays; u,v,w updated in place.
// markers/A/b/x/div outputs may be null.
int orc_pressure_solve_f32(const orc_grid* gr, float dt, float density, float* u, float* v,
                           float* w, const float* us, const float* vs, const float* ws,
                           const float* solid, const float* fluid, int32_t tol_mode,
                           double tol, int64_t max_it, int32_t jacobi, float* markers,
                           float* A, float* b, float* x, orc_stats* st) {
    Grid<float> g = to_grid<float>(gr);
    PressureOut<float> o = pressure_solve(g, dt, density, u, v, w, us, vs, ws, solid, fluid,
                                          tol_mode, tol, max_it, jacobi != 0);
    if (markers) std::copy(o.markers.begin(), o.markers.end(), markers);
    if (A) std::copy(o.A.begin(), o.A.end(), A);
    if (b) std::copy(o.b.begin(), o.b.end(), b);
    if (x) std::copy(o.x.begin(), o.x.end(), x);
    if (st) {
        st->cg_iterations = o.cg.iterations;
        st->converged = o.cg.converged;
        st->residual_recurrence = o.cg.r_norm;
        st->b_norm = o.cg.b_norm;
        st->residual_rel = true_residual_rel(g, o.A.data(), o.b.data(), o.x.data());
    }
    return 0;
}

// Same in FP64 with float inputs widened; outputs in double.
int orc_pressure_solve_f64(const orc_grid* gr, float dt, float density, const float* u,
                           const float* v, const float* w, const float* us, const float* vs,
                           const float* ws, const float* solid, const float* fluid,
                           int32_t tol_mode, double tol, int64_t max_it, int32_t jacobi,
                           double* uo, double* vo, double* wo, double* x, orc_stats* st) {
    Grid<double> g = to_grid<double>(gr);
    std::vector<double> U = vec<double>(u, nfx(gr)), V = vec<double>(v, nfy(gr)),
                        W = vec<double>(w, nfz(gr));
    std::vector<double> US = vec<double>(us, nfx(gr)), VS = vec<double>(vs, nfy(gr)),
                        WS = vec<double>(ws, nfz(gr));
    std::vector<double> S = vec<double>(solid, ncc(gr)), F = vec<double>(fluid, ncc(gr));
    PressureOut<double> o = pressure_solve(g, double(dt), double(density), U.data(), V.data(),
                                           W.data(), US.data(), VS.data(), WS.data(), S.data(),
                                           F.data(), tol_mode, tol, max_it, jacobi != 0);
    if (uo) put_d(U, uo);
    if (vo) put_d(V, vo);
    if (wo) put_d(W, wo);
    if (x) put_d(o.x, x);
    if (st) {
        st->cg_iterations = o.cg.iterations;
        st->converged = o.cg.converged;
        st->residual_recurrence = o.cg.r_norm;
        st->b_norm = o.cg.b_norm;
        st->residual_rel = true_residual_rel(g, o.A.data(), o.b.data(), o.x.data());
    }
    return 0;
}

// ---- scene ---------------------------------------------------------------
void* orc_scene_create(const orc_grid* gr, int32_t f64) {
    auto* h = new SceneHandle;
    h->f64 = f64;
    if (f64) h->s64.reset(new Scene<double>(to_grid<double>(gr)));
    else h->s32.reset(new Scene<float>(to_grid<float>(gr)));
    return h;
}

void orc_scene_destroy(void* hp) { delete static_cast<SceneHandle*>(hp); }

int orc_scene_add_solid(void* hp, const float* sdf, int64_t n, const float vel[3]) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        using R = typename std::decay_t<decltype(s.u)>::value_type;
        if (size_t(n) != s.g.cells()) return 1;
        s.sdf0.push_back(vec<R>(sdf, size_t(n)));
        for (int k = 0; k < 3; ++k) { s.svel.push_back(R(vel[k])); s.strans.push_back(R(0)); }
        s.boundary_from_solids = true;
        return 0;
    });
}

int orc_scene_fluid_box(void* hp, const float mn[3], const float mx[3]) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        using R = typename std::decay_t<decltype(s.u)>::value_type;
        R a[3] = {R(mn[0]), R(mn[1]), R(mn[2])}, b[3] = {R(mx[0]), R(mx[1]), R(mx[2])};
        fluid_domain_box(s.g, a, b, s.phi.data());
        return 0;
    });
}

// Uploading the solid marker or solid velocity replaces what the solids
// produce, as in CPressureSolver::solvePressure called with caller fields.
int orc_scene_upload(void* hp, int32_t field, const float* data, int64_t n) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        auto* f = field_ptr(s, field);
        if (!f || f->size() != size_t(n)) return 1;
        using R = typename std::decay_t<decltype(*f)>::value_type;
        for (int64_t i = 0; i < n; ++i) (*f)[i] = R(data[i]);
        if (field >= 4 && field <= 7) s.boundary_from_solids = false;
        return 0;
    });
}

int orc_scene_download(void* hp, int32_t field, float* data, int64_t n) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        auto* f = field_ptr(s, field);
        if (!f || f->size() != size_t(n)) return 1;
        put(*f, data);
        return 0;
    });
}

int orc_scene_download_f64(void* hp, int32_t field, double* data, int64_t n) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        auto* f = field_ptr(s, field);
        if (!f || f->size() != size_t(n)) return 1;
        put_d(*f, data);
        return 0;
    });
}

int orc_scene_update_boundaries(void* hp, const orc_params* p) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        using R = typename std::decay_t<decltype(s.u)>::value_type;
        s.update_boundaries(R(p->dt));
        return 0;
    });
}
int orc_scene_advect(void* hp, const orc_params* p) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) { s.advect(*p); return 0; });
}
int orc_scene_apply_forces(void* hp, const orc_params* p) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) { s.forces(*p); return 0; });
}
int orc_scene_pressure_solve(void* hp, const orc_params* p, orc_stats* st) {
    auto* h = static_cast<SceneHandle*>(hp);
    if (st) { *st = orc_stats{}; st->converged = 1; }
    return with_scene(h, [&](auto& s) { s.pressure(*p, st); return 0; });
}
int orc_scene_extrapolate(void* hp, const orc_params* p) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) { return s.extrapolate(*p); });
}
int orc_scene_step(void* hp, const orc_params* p, int32_t nsteps, orc_stats* st) {
    auto* h = static_cast<SceneHandle*>(hp);
    if (st) { *st = orc_stats{}; st->converged = 1; }
    return with_scene(h, [&](auto& s) {
        for (int i = 0; i < nsteps; ++i) s.step(*p, st);
        return 0;
    });
}

// Translation of solid `idx` (3 values) and its last re-sampled SDF.
int orc_scene_solid_state(void* hp, int32_t idx, float* translation, float* sdf, int64_t n) {
    auto* h = static_cast<SceneHandle*>(hp);
    return with_scene(h, [&](auto& s) {
        if (idx < 0 || size_t(idx) >= s.sdf0.size()) return 1;
        for (int k = 0; k < 3; ++k) translation[k] = float(s.strans[3 * idx + k]);
        if (sdf) {
            if (size_t(idx) >= s.sdf_cur.size() || size_t(n) != s.g.cells()) return 1;
            put(s.sdf_cur[idx], sdf);
        }
        return 0;
    });
}

// ---- particle sets and particle solvers (FP32) -------------------------------
uint64_t orc_splitmix64(uint64_t z) { return splitmix64(z); }

int64_t orc_generate_particles(const orc_grid* gr, const float* fluid, const float* solid, int32_t per_cell,
                               uint64_t seed, float* out, int64_t cap) {
    const std::vector<float> p = generate_particles(to_grid<float>(gr), fluid, solid, per_cell, seed);
    const int64_t n = int64_t(p.size() / 3);
    for (int64_t i = 0; i < 3 * std::min(n, cap); ++i) out[i] = p[size_t(i)];
    return n;
}

void* orc_psolver_create(const orc_grid* gr, int32_t kind, int32_t alg, float cfl, float mixing, int64_t n,
                         const float* pos) {
    auto* ps = new ParticleSolver<float>;
    ps->g = to_grid<float>(gr);
    ps->kind = kind; ps->alg = alg; ps->cfl = cfl; ps->mixing = mixing;
    ps->pos.assign(pos, pos + 3 * n);
    ps->init(size_t(n));
    return ps;
}

void orc_psolver_destroy(void* h) { delete static_cast<ParticleSolver<float>*>(h); }

int orc_psolver_advect_velocity(void* h, const float* iu, const float* iv, const float* iw, const float* vu,
                                const float* vv, const float* vw, float dt, int32_t order, int32_t fix, float* ou,
                                float* ov, float* ow) {
    auto* ps = static_cast<ParticleSolver<float>*>(h);
    const float* in[3] = {iu, iv, iw};
    const float* V[3] = {vu, vv, vw};
    float* out[3] = {ou, ov, ow};
    ps->advect_velocity(in, V, dt, order, fix != 0, out);
    return 0;
}

int orc_psolver_advect_scalar(void* h, const float* in, float* out) {
    static_cast<ParticleSolver<float>*>(h)->advect_scalar(in, out);
    return 0;
}

int orc_psolver_advect_vector(void* h, const float* ix, const float* iy, const float* iz, float* ox, float* oy,
                              float* oz) {
    const float* in[3] = {ix, iy, iz};
    float* out[3] = {ox, oy, oz};
    static_cast<ParticleSolver<float>*>(h)->advect_vector(in, out);
    return 0;
}

// which: 0 positions, 1 velocities, 3 scalar values, 4 vector values (the
// numbering of MGF_PARTICLE_*).
int orc_psolver_get(void* h, int32_t which, float* out, int64_t count) {
    auto* ps = static_cast<ParticleSolver<float>*>(h);
    const std::vector<float>* v = which == 0 ? &ps->pos : which == 1 ? &ps->vel : which == 3 ? &ps->scalar
                                : which == 4 ? &ps->vec3 : nullptr;
    if (!v || int64_t(v->size()) != count) return 1;
    std::copy(v->begin(), v->end(), out);
    return 0;
}

int orc_scene_step_particles(void* hp, void* psh, const orc_params* p, int32_t nsteps, orc_stats* st) {
    auto* h = static_cast<SceneHandle*>(hp);
    if (h->f64) return 1;
    if (st) { *st = orc_stats{}; st->converged = 1; }
    auto* ps = static_cast<ParticleSolver<float>*>(psh);
    for (int i = 0; i < nsteps; ++i) h->s32->step_particles(*ps, *p, st);
    return 0;
}

int orc_scene_pressure_solve_particles(void* hp, void* psh, const orc_params* p, orc_stats* st) {
    auto* h = static_cast<SceneHandle*>(hp);
    if (h->f64) return 1;
    if (st) { *st = orc_stats{}; st->converged = 1; }
    h->s32->pressure_particles(*static_cast<ParticleSolver<float>*>(psh), *p, st);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
