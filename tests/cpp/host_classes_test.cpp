// The reference-shaped C++ classes (hybrid-fluid-simulation_amd/host/mgf_fluid.hpp)
// against the C ABI they wrap: every result must be bit-identical to the same
// work done through mgf.h directly.  Needs an MI355X; exits non-zero on a
// failure.  Built in-tree by hybrid-fluid-simulation_amd/build.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mgf_fluid.hpp"

using namespace mgf;

static int g_checks = 0, g_failures = 0;

#define EXPECT(cond, ...)                                          \
    do {                                                           \
        ++g_checks;                                                \
        if (!(cond)) {                                             \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);       \
            std::printf(__VA_ARGS__);                              \
            std::printf("\n");                                     \
            ++g_failures;                                          \
        }                                                          \
    } while (0)

static bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

static std::vector<float> host_field(mgf_scene s, int field) {
    int64_t n = 0;
    check(mgf_scene_field_count(s, field, &n), "count");
    std::vector<float> h(static_cast<size_t>(n));
    check(mgf_scene_download(s, field, h.data(), n), "download");
    return h;
}

// Dam break on [0,1]^3: n interior cells per axis, one ghost layer of solid
// wall, fluid block x < 0.4, y < 0.6 (all z).
struct Dam {
    int n;
    Vector3i res;
    Vector3 o, h;
    Real dt;
    std::vector<float> wall;   // distance to the walls, < 0 on the ghost layer
    std::vector<float> colour[3];

    explicit Dam(int n_) : n(n_) {
        const float hh = 1.0f / float(n);
        res = Vector3i(n + 2, n + 2, n + 2);
        o = Vector3(-hh, -hh, -hh);
        h = Vector3(hh, hh, hh);
        dt = 2e-3f;
        const int g = n + 2;
        wall.resize(size_t(g) * g * g);
        for (int k = 0; k < 3; ++k) colour[k].resize(wall.size());
        for (int z = 0; z < g; ++z)
            for (int y = 0; y < g; ++y)
                for (int x = 0; x < g; ++x) {
                    const double c[3] = {double(o.x + 0.5f * hh + hh * float(x)),
                                         double(o.y + 0.5f * hh + hh * float(y)),
                                         double(o.z + 0.5f * hh + hh * float(z))};
                    double d = 1e30;
                    for (int a = 0; a < 3; ++a) d = std::min(d, std::min(c[a], 1.0 - c[a]));
                    const size_t i = (size_t(z) * g + y) * g + x;
                    wall[i] = float(d);
                    for (int k = 0; k < 3; ++k) colour[k][i] = float(c[k]);
                }
    }
    CCellCenteredScalarField wallField() const { return CCellCenteredScalarField(res, o, h, wall.data()); }
    CCellCenteredVectorField colourField() const {
        return CCellCenteredVectorField(res, o, h, colour[0].data(), colour[1].data(), colour[2].data());
    }
    mgf_scene_desc desc(const PressureOptions& p) const {
        mgf_scene_desc d = desc_for(dt, res, o, h);
        apply_options(d, p);
        return d;
    }
};

static PressureOptions mixed_mg() {
    PressureOptions p;
    p.precision = MGF_PRECISION_MIXED;
    p.preconditioner = MGF_PRECOND_MULTIGRID;
    p.tol_mode = MGF_TOL_RELATIVE;
    p.tolerance = 1e-6;
    return p;
}

static void compare_scene(mgf_scene a, mgf_scene b, const char* what) {
    const int fields[] = {MGF_FIELD_VEL_X, MGF_FIELD_VEL_Y, MGF_FIELD_VEL_Z, MGF_FIELD_FLUID_DOMAIN,
                          MGF_FIELD_COLOR_X, MGF_FIELD_COLOR_Y, MGF_FIELD_COLOR_Z, MGF_FIELD_PRESSURE};
    for (int f : fields) EXPECT(same_bits(host_field(a, f), host_field(b, f)), "%s: field %d differs", what, f);
}

// A CAdvectionSolver of the caller's own (the virtual strategy slot): it
// forwards to CSemiLagrangian and counts the calls.
class CountingAdvection : public CAdvectionSolver {
public:
    explicit CountingAdvection(const Dam& d) : m_SL(d.res, d.o, d.h) {}
    void advect(const CCellCenteredScalarField& i, const CFaceCenteredVectorField& v, Real dt,
                CCellCenteredScalarField& o, EAdvectionAccuracy a, const CCellCenteredScalarField& s) override {
        ++calls;
        m_SL.advect(i, v, dt, o, a, s);
    }
    void advect(const CCellCenteredVectorField& i, const CFaceCenteredVectorField& v, Real dt,
                CCellCenteredVectorField& o, EAdvectionAccuracy a, const CCellCenteredScalarField& s) override {
        ++calls;
        m_SL.advect(i, v, dt, o, a, s);
    }
    void advect(const CFaceCenteredVectorField& i, const CFaceCenteredVectorField& v, Real dt,
                CFaceCenteredVectorField& o, EAdvectionAccuracy a, const CCellCenteredScalarField& s) override {
        ++calls;
        m_SL.advect(i, v, dt, o, a, s);
    }
    int calls = 0;

private:
    CSemiLagrangian m_SL;
};

static void build_solver(CGridFluidSolver& s, const Dam& d, bool colour) {
    s.resizeGridFluidSolver(d.dt, d.res, d.o, d.h);
    s.setPressureOptions(mixed_mg());
    s.addSolidBoundary(d.wallField());
    s.generateFluidDomainFromBBox(Vector3(0, 0, 0), Vector3(0.4f, 0.6f, 1.0f));
    if (colour) s.setColorField(d.colourField());
}

static SceneHandle build_scene(const Dam& d, bool colour, const Vector3& force = Vector3(0, 0, 0)) {
    mgf_scene_desc desc = d.desc(mixed_mg());
    desc.external_force[0] = force.x; desc.external_force[1] = force.y; desc.external_force[2] = force.z;
    SceneHandle sc(desc);
    const float v0[3] = {0, 0, 0};
    check(mgf_scene_add_solid(sc.get(), d.wall.data(), int64_t(d.wall.size()), v0), "add_solid");
    const float mn[3] = {0, 0, 0}, mx[3] = {0.4f, 0.6f, 1.0f};
    check(mgf_scene_fluid_box(sc.get(), mn, mx), "fluid_box");
    if (colour)
        for (int k = 0; k < 3; ++k)
            check(mgf_scene_upload(sc.get(), MGF_FIELD_COLOR_X + k, d.colour[k].data(), int64_t(d.wall.size())),
                  "upload colour");
    return sc;
}

int main() {
    const Dam d(20);
    const int steps = 3;

    // 1. CGridFluidSolver::update (library step) == mgf_step on the same scene.
    CGridFluidSolver fast;
    build_solver(fast, d, true);
    for (int i = 0; i < steps; ++i) fast.update();
    SceneHandle ref = build_scene(d, true);
    mgf_step_stats st;
    check(mgf_step(ref.get(), steps, &st), "mgf_step");
    compare_scene(fast.scene(), ref.get(), "update() vs mgf_step");
    EXPECT(fast.getCurSimulationTime() == d.dt + d.dt + d.dt, "simulation time %g", double(fast.getCurSimulationTime()));
    EXPECT(st.converged == 1 && st.residual_rel <= 1e-6, "reference scene converged %d resid %g", st.converged,
           st.residual_rel);

    // 2. A caller's CAdvectionSolver through the virtual slot gives the same
    //    fields (velocity, colour and domain advected with the old velocity).
    CGridFluidSolver generic;
    build_solver(generic, d, true);
    auto counting = std::make_shared<CountingAdvection>(d);
    generic.setAdvectionSolver(counting);
    for (int i = 0; i < steps; ++i) generic.update();
    compare_scene(generic.scene(), ref.get(), "virtual advection slot vs mgf_step");
    EXPECT(counting->calls == 3 * steps, "advect calls %d", counting->calls);

    // 3. updateWithoutPressure + solvePressure == update, with an external force.
    CGridFluidSolver split;
    build_solver(split, d, false);
    split.addExternalForce(Vector3(0.5f, 0, -0.25f));
    for (int i = 0; i < steps; ++i) {
        split.updateWithoutPressure();
        split.solvePressure();
    }
    SceneHandle ref_f = build_scene(d, false, Vector3(0.5f, 0, -0.25f));
    check(mgf_step(ref_f.get(), steps, &st), "mgf_step");
    compare_scene(split.scene(), ref_f.get(), "split step vs mgf_step");

    // 4. CPressureSolver::solvePressure on caller fields == the scene's solve.
    SceneHandle pre = build_scene(d, false);
    check(mgf_scene_update_boundaries(pre.get()), "update_boundaries");
    check(mgf_scene_advect(pre.get()), "advect");
    check(mgf_scene_apply_forces(pre.get()), "apply_forces");
    CFaceCenteredVectorField vel(d.res, d.o, d.h), svel(d.res, d.o, d.h);
    CCellCenteredScalarField ssdf(d.res, d.o, d.h), phi(d.res, d.o, d.h);
    for (int k = 0; k < 3; ++k) {
        download(pre.get(), MGF_FIELD_VEL_X + k, vel.component(k), vel.faces(k));
        download(pre.get(), MGF_FIELD_SOLID_VEL_X + k, svel.component(k), svel.faces(k));
    }
    download(pre.get(), MGF_FIELD_SOLID_SDF, ssdf.getGridDataGPUPtr(), ssdf.getDataSize());
    download(pre.get(), MGF_FIELD_FLUID_DOMAIN, phi.getGridDataGPUPtr(), phi.getDataSize());
    CPressureSolver ps(d.res, d.o, d.h);
    ps.setOptions(mixed_mg());
    ps.solvePressure(vel, d.dt, svel, ssdf, phi);
    check(mgf_scene_pressure_solve(pre.get(), &st), "pressure_solve");
    for (int k = 0; k < 3; ++k)
        EXPECT(same_bits(vel.getHostData(k), host_field(pre.get(), MGF_FIELD_VEL_X + k)),
               "CPressureSolver velocity component %d", k);
    EXPECT(same_bits(ps.getConstVectorx().getHostData(), host_field(pre.get(), MGF_FIELD_PRESSURE)),
           "CPressureSolver pressure");
    EXPECT(same_bits(ps.getMarkers().getHostData(), host_field(pre.get(), MGF_FIELD_MARKERS)), "markers");
    EXPECT(ps.getLastStats().converged == 1, "CPressureSolver converged");

    //    The reference's default (plain FP32, sqrt(||r||) < 1e-3) also runs.
    CPressureSolver plain(d.res, d.o, d.h);
    CFaceCenteredVectorField vel2 = vel;
    plain.solvePressure(vel2, d.dt, svel, ssdf, phi);
    EXPECT(plain.getLastStats().cg_iterations > 0, "plain CG ran");

    // 5. CExternalForcesSolver::applyExternalForces == mgf_scene_apply_forces.
    SceneHandle frc = build_scene(d, false, Vector3(1.0f, 2.0f, 0));
    CFaceCenteredVectorField v3(d.res, d.o, d.h);
    for (int k = 0; k < 3; ++k) download(frc.get(), MGF_FIELD_VEL_X + k, v3.component(k), v3.faces(k));
    CExternalForcesSolver ef(d.res, d.o, d.h);
    ef.addExternalForces(Vector3(1.0f, 2.0f, 0));
    ef.applyExternalForces(v3, d.dt);
    check(mgf_scene_apply_forces(frc.get()), "apply_forces");
    for (int k = 0; k < 3; ++k)
        EXPECT(same_bits(v3.getHostData(k), host_field(frc.get(), MGF_FIELD_VEL_X + k)), "forces component %d", k);

    // 6. CSemiLagrangian: zero velocity back-traces to the same points;
    //    advect(FCV) is mgf_op_advect_fc.
    CSemiLagrangian sl(d.res, d.o, d.h);
    CFaceCenteredVectorField zero(d.res, d.o, d.h);
    std::vector<float> px(d.wall.size()), py(px.size()), pz(px.size());
    for (size_t i = 0; i < px.size(); ++i) {
        px[i] = 0.01f * float(i % 97);
        py[i] = 0.02f * float(i % 31);
        pz[i] = 0.03f * float(i % 13);
    }
    CCellCenteredVectorField pin(d.res, d.o, d.h, px.data(), py.data(), pz.data()), pout;
    sl.backTrace(pin, zero, d.dt, pout, EAdvectionAccuracy::RK3);
    EXPECT(same_bits(pout.getHostData(0), px) && same_bits(pout.getHostData(1), py) &&
               same_bits(pout.getHostData(2), pz),
           "zero-velocity back-trace moved points");
    CFaceCenteredVectorField adv;
    sl.advect(vel, vel, d.dt, adv, EAdvectionAccuracy::RK2);
    CFaceCenteredVectorField adv2(d.res, d.o, d.h);
    const mgf_grid g = vel.toMgfGrid();
    check(mgf_op_advect_fc(ctx(), &g, vel.getConstGridDataXGPUPtr(), vel.getConstGridDataYGPUPtr(),
                           vel.getConstGridDataZGPUPtr(), vel.getConstGridDataXGPUPtr(), vel.getConstGridDataYGPUPtr(),
                           vel.getConstGridDataZGPUPtr(), d.dt, MGF_RK2, adv2.getGridDataXGPUPtr(),
                           adv2.getGridDataYGPUPtr(), adv2.getGridDataZGPUPtr()),
          "advect_fc");
    for (int k = 0; k < 3; ++k) EXPECT(same_bits(adv.getHostData(k), adv2.getHostData(k)), "advect FCV %d", k);

    // 7. generateCurFluidSDF / generateCurFluidDensity read the library's fields.
    fast.generateCurFluidSDF();
    fast.generateCurFluidDensity();
    EXPECT(same_bits(fast.getFluidSDFField().getHostData(), host_field(ref.get(), MGF_FIELD_FLUID_SDF)),
           "fluid SDF");
    EXPECT(same_bits(fast.getFluidDensityField().getHostData(), host_field(ref.get(), MGF_FIELD_FLUID_DENSITY)),
           "fluid density");

    // 8. Field copies are deep.
    CCellCenteredScalarField a = d.wallField(), b = a;
    std::vector<float> ones(d.wall.size(), 1.0f);
    a.resize(d.res, d.o, d.h, ones.data());
    EXPECT(same_bits(b.getHostData(), d.wall), "copy is not deep");

    // 9. Particle advection solvers in the strategy slot: the step with
    //    particle markers (mgf_step_particles) and the split path through the
    //    virtual advect both follow the library's own particle step.  The
    //    particle-to-grid sums are FP32 atomics, so runs agree to rounding.
    {
        CGridFluidSolver pf;
        build_solver(pf, d, true);
        const CCellCenteredScalarField wall = d.wallField();
        const CCellCenteredScalarField dom = pf.getFluidDomainField();
        auto flip = std::make_shared<CFluidImplicitParticle>();
        flip->setSeed(77);
        flip->resizeFluidImplicitParticle(dom, wall, 4, d.res, d.o, d.h, 1.0f);
        pf.setAdvectionSolver(flip);
        for (int i = 0; i < steps; ++i) pf.update();

        SceneHandle pref = build_scene(d, true);
        mgf_particles parts = nullptr;
        const mgf_grid gg = wall.toMgfGrid();
        check(mgf_particles_generate(ctx(), &gg, dom.getConstGridDataGPUPtr(), wall.getConstGridDataGPUPtr(), 4, 77,
                                     &parts), "generate");
        EXPECT(flip->getEulerParticles().getNumOfParticles() > 0, "no particles generated");
        mgf_psolver solver = nullptr;
        check(mgf_psolver_create(ctx(), &gg, MGF_PSOLVER_FLIP, parts, 1.0f, MGF_PG_LINEAR, 0.1f, &solver), "psolver");
        check(mgf_step_particles(pref.get(), solver, steps, &st), "mgf_step_particles");
        const int vf[] = {MGF_FIELD_VEL_X, MGF_FIELD_VEL_Y, MGF_FIELD_VEL_Z, MGF_FIELD_COLOR_X};
        for (int f : vf) {
            const std::vector<float> a = host_field(pf.scene(), f), b = host_field(pref.get(), f);
            double num = 0, den = 0;
            for (size_t i = 0; i < a.size(); ++i) {
                num += double(a[i] - b[i]) * double(a[i] - b[i]);
                den += double(b[i]) * double(b[i]);
            }
            EXPECT(num <= 1e-8 * std::max(den, 1e-30), "particle step field %d: rel L2 %g", f,
                   std::sqrt(num / std::max(den, 1e-30)));
        }
        std::vector<float> pa = flip->getEulerParticles().getParticlesPos(), pb(pa.size());
        check(mgf_particles_download(parts, MGF_PARTICLE_POS, pb.data(), int64_t(pb.size())), "download");
        float worst = 0;
        for (size_t i = 0; i < pa.size(); ++i) worst = std::max(worst, std::fabs(pa[i] - pb[i]));
        EXPECT(worst <= 1e-4f, "particle positions differ by %g", double(worst));
        check(mgf_psolver_destroy(solver), "destroy");
        check(mgf_particles_destroy(parts), "destroy");

        // the split path: virtual advect on copies + pressure with particle markers
        CGridFluidSolver ps2;
        build_solver(ps2, d, true);
        auto pic = std::make_shared<CParticleInCell>(dom, wall, 2, d.res, d.o, d.h, 1.0f);
        ps2.setAdvectionSolver(pic);
        CGridFluidSolver ps3;
        build_solver(ps3, d, true);
        auto pic3 = std::make_shared<CParticleInCell>(dom, wall, 2, d.res, d.o, d.h, 1.0f);
        ps3.setAdvectionSolver(pic3);
        for (int i = 0; i < 2; ++i) {
            ps2.updateWithoutPressure();
            ps2.solvePressure();
            ps3.update();
        }
        for (int f : vf) {
            const std::vector<float> a = host_field(ps2.scene(), f), b = host_field(ps3.scene(), f);
            double num = 0, den = 0;
            for (size_t i = 0; i < a.size(); ++i) {
                num += double(a[i] - b[i]) * double(a[i] - b[i]);
                den += double(b[i]) * double(b[i]);
            }
            EXPECT(num <= 1e-8 * std::max(den, 1e-30), "split particle step field %d: rel L2 %g", f,
                   std::sqrt(num / std::max(den, 1e-30)));
        }

        // CEulerParticles alone: generation and the per-cell count
        CEulerParticles ep;
        ep.generateParticlesInFluid(dom, wall, 3, 5);
        CCellCenteredScalarField count(d.res, d.o, d.h);
        ep.statisticalFluidDensity(count);
        double total = 0;
        for (float c : count.getHostData()) total += c;
        EXPECT(total == double(ep.getNumOfParticles()), "counted %g of %u particles", total, ep.getNumOfParticles());
        std::vector<float> vel0(3 * size_t(ep.getNumOfParticles()), 0.5f);
        ep.setParticlesVel(vel0.data());
        EXPECT(same_bits(ep.getParticlesVel(), vel0), "particle velocities round trip");
    }

    std::printf("host classes: %d checks, %d failures\n", g_checks, g_failures);
    return g_failures ? 1 : 0;
}
