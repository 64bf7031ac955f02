// mgf_fluid.hpp — the reference's grid-step classes over the C ABI of mgf.h.
//
// Class names, method names, argument order and defaults follow the
// reference's headers, so code written against them compiles against these:
//   CGridFluidSolver          MPhysics/GridFluidSolver.h:17-138
//   CPressureSolver           MPhysics/PressureSolver.h:18-51
//   CAdvectionSolver          MPhysics/AdvectionSolver.h:12-40
//   CSemiLagrangian           MPhysics/SemiLagrangian.h:4-50
//   CExternalForcesSolver     MPhysics/ExternalForcesSolver.h:6-28
//   CEulerParticles           MPhysics/EulerParticles.h:5-64
//   CParticleInCell, CFluidImplicitParticle, CMixPICAndFLIP
//                             MPhysics/ParticleInCell.h, FluidImplicitParticle.h, MixPICAndFLIP.h
//   CCellCenteredScalarField, CCellCenteredVectorField, CFaceCenteredVectorField
//                             (MPhysics/CellCenteredScalarField.h, CellCenteredVectorField.h,
//                             FaceCenteredVectorField.h; layouts Grid.h:5-24, VectorGrid.h:34-36)
// Fields live in device memory the class owns (allocated through the
// library); copies are deep device copies, as the reference's operator=.
// Every call goes through libmgf.so; a failing call throws mgf::Error (the C
// ABI itself never throws).
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mgf.h"

namespace mgf {

using Real = float;

struct Vector3 {
    Real x = 0, y = 0, z = 0;
    Vector3() = default;
    Vector3(Real a, Real b, Real c) : x(a), y(b), z(c) {}
    Vector3 operator+(const Vector3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector3 operator*(Real s) const { return {x * s, y * s, z * s}; }
    Vector3& operator+=(const Vector3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    bool operator==(const Vector3& o) const { return x == o.x && y == o.y && z == o.z; }
};

struct Vector3i {
    int x = 0, y = 0, z = 0;
    Vector3i() = default;
    Vector3i(int a, int b, int c) : x(a), y(b), z(c) {}
    Vector3i operator+(const Vector3i& o) const { return {x + o.x, y + o.y, z + o.z}; }
    bool operator==(const Vector3i& o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const Vector3i& o) const { return !(*this == o); }
};

// EAdvectionAccuracy (MPhysics/SemiLagrangian.cpp:188-347)
enum class EAdvectionAccuracy { RK1 = MGF_RK1, RK2 = MGF_RK2, RK3 = MGF_RK3 };

struct Error : std::runtime_error {
    mgf_status status;
    Error(mgf_status s, const std::string& where)
        : std::runtime_error(where + ": " + mgf_last_error() + " (" + mgf_status_string(s) + ")"), status(s) {}
};

inline void check(mgf_status s, const char* where) {
    if (s != MGF_OK) throw Error(s, where);
}

// CCudaContextManager (MPhysics/CudaContextManager.cpp:3-40): one context on
// device 0 for the process.
class CCudaContextManager {
public:
    static CCudaContextManager& getInstance() {
        static CCudaContextManager inst;
        return inst;
    }
    mgf_context context() const { return m_Ctx; }
    ~CCudaContextManager() {
        if (m_Ctx) mgf_shutdown(m_Ctx);
    }
    CCudaContextManager(const CCudaContextManager&) = delete;
    CCudaContextManager& operator=(const CCudaContextManager&) = delete;

private:
    CCudaContextManager() { check(mgf_init(0, &m_Ctx), "mgf_init"); }
    mgf_context m_Ctx = nullptr;
};

inline mgf_context ctx() { return CCudaContextManager::getInstance().context(); }

// Device array of Real owned through the library.
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray& o) { *this = o; }
    DeviceArray& operator=(const DeviceArray& o) {
        if (this == &o) return *this;
        resize(o.m_N);
        if (m_N) check(mgf_memcpy_d2d(ctx(), m_P, o.m_P, int64_t(m_N * sizeof(Real))), "mgf_memcpy_d2d");
        return *this;
    }
    DeviceArray(DeviceArray&& o) noexcept : m_P(o.m_P), m_N(o.m_N) { o.m_P = nullptr; o.m_N = 0; }
    ~DeviceArray() { release(); }

    void resize(size_t n) {
        if (n == m_N) return;
        release();
        if (n) {
            void* p = nullptr;
            check(mgf_device_alloc(ctx(), int64_t(n * sizeof(Real)), &p), "mgf_device_alloc");
            m_P = static_cast<Real*>(p);
            m_N = n;
        }
    }
    void fromHost(const Real* h) {
        if (m_N) check(mgf_memcpy_h2d(ctx(), m_P, h, int64_t(m_N * sizeof(Real))), "mgf_memcpy_h2d");
    }
    void zero() {
        std::vector<Real> z(m_N, Real(0));
        fromHost(z.data());
    }
    std::vector<Real> toHost() const {
        std::vector<Real> h(m_N);
        if (m_N) check(mgf_memcpy_d2h(ctx(), h.data(), m_P, int64_t(m_N * sizeof(Real))), "mgf_memcpy_d2h");
        return h;
    }
    Real* data() { return m_P; }
    const Real* data() const { return m_P; }
    size_t size() const { return m_N; }

private:
    void release() {
        if (m_P) mgf_device_free(ctx(), m_P);
        m_P = nullptr;
        m_N = 0;
    }
    Real* m_P = nullptr;
    size_t m_N = 0;
};

// CGrid (MPhysics/Grid.h:5-24): resolution, origin, spacing.
class CGrid {
public:
    const Vector3i& getResolution() const { return m_Resolution; }
    const Vector3& getOrigin() const { return m_Origin; }
    const Vector3& getSpacing() const { return m_Spacing; }
    mgf_grid toMgfGrid() const {
        mgf_grid g;
        g.nx = m_Resolution.x; g.ny = m_Resolution.y; g.nz = m_Resolution.z;
        g.origin[0] = m_Origin.x; g.origin[1] = m_Origin.y; g.origin[2] = m_Origin.z;
        g.spacing[0] = m_Spacing.x; g.spacing[1] = m_Spacing.y; g.spacing[2] = m_Spacing.z;
        return g;
    }

protected:
    void setGrid(const Vector3i& r, const Vector3& o, const Vector3& h) {
        m_Resolution = r; m_Origin = o; m_Spacing = h;
    }
    size_t cells() const { return size_t(m_Resolution.x) * m_Resolution.y * m_Resolution.z; }
    Vector3i m_Resolution;
    Vector3 m_Origin;
    Vector3 m_Spacing = Vector3(1, 1, 1);
};

// CCellCenteredScalarField (resize zero-fills or copies from the host,
// MPhysics/ScalarGrid.cpp:33-47).
class CCellCenteredScalarField : public CGrid {
public:
    CCellCenteredScalarField() = default;
    CCellCenteredScalarField(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                             const Vector3& h = Vector3(1, 1, 1), const Real* cpu = nullptr) {
        resize(r, o, h, cpu);
    }
    void resize(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0), const Vector3& h = Vector3(1, 1, 1),
                const Real* cpu = nullptr) {
        setGrid(r, o, h);
        m_Data.resize(cells());
        if (cpu) m_Data.fromHost(cpu); else m_Data.zero();
    }
    const Real* getConstGridDataGPUPtr() const { return m_Data.data(); }
    Real* getGridDataGPUPtr() { return m_Data.data(); }
    std::vector<Real> getHostData() const { return m_Data.toHost(); }
    size_t getDataSize() const { return m_Data.size(); }
    bool empty() const { return m_Data.size() == 0; }

private:
    DeviceArray m_Data;
};

// CCellCenteredVectorField: three SoA components of nx*ny*nz each.
class CCellCenteredVectorField : public CGrid {
public:
    CCellCenteredVectorField() = default;
    CCellCenteredVectorField(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                             const Vector3& h = Vector3(1, 1, 1), const Real* x = nullptr,
                             const Real* y = nullptr, const Real* z = nullptr) {
        resize(r, o, h, x, y, z);
    }
    void resize(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0), const Vector3& h = Vector3(1, 1, 1),
                const Real* x = nullptr, const Real* y = nullptr, const Real* z = nullptr) {
        setGrid(r, o, h);
        const Real* src[3] = {x, y, z};
        for (int k = 0; k < 3; ++k) {
            m_Data[k].resize(cells());
            if (src[k]) m_Data[k].fromHost(src[k]); else m_Data[k].zero();
        }
    }
    const Real* getConstGridDataXGPUPtr() const { return m_Data[0].data(); }
    const Real* getConstGridDataYGPUPtr() const { return m_Data[1].data(); }
    const Real* getConstGridDataZGPUPtr() const { return m_Data[2].data(); }
    Real* getGridDataXGPUPtr() { return m_Data[0].data(); }
    Real* getGridDataYGPUPtr() { return m_Data[1].data(); }
    Real* getGridDataZGPUPtr() { return m_Data[2].data(); }
    const Real* component(int k) const { return m_Data[k].data(); }
    Real* component(int k) { return m_Data[k].data(); }
    std::vector<Real> getHostData(int k) const { return m_Data[k].toHost(); }
    size_t getDataSize() const { return m_Data[0].size(); }
    bool empty() const { return m_Data[0].size() == 0; }

private:
    DeviceArray m_Data[3];
};

// CFaceCenteredVectorField (MPhysics/FaceCenteredVectorGrid.cpp:74-123):
// extents (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1).
class CFaceCenteredVectorField : public CGrid {
public:
    CFaceCenteredVectorField() = default;
    CFaceCenteredVectorField(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                             const Vector3& h = Vector3(1, 1, 1), const Real* x = nullptr,
                             const Real* y = nullptr, const Real* z = nullptr) {
        resize(r, o, h, x, y, z);
    }
    void resize(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0), const Vector3& h = Vector3(1, 1, 1),
                const Real* x = nullptr, const Real* y = nullptr, const Real* z = nullptr) {
        setGrid(r, o, h);
        const Real* src[3] = {x, y, z};
        for (int k = 0; k < 3; ++k) {
            m_Data[k].resize(faces(k));
            if (src[k]) m_Data[k].fromHost(src[k]); else m_Data[k].zero();
        }
    }
    size_t faces(int k) const {
        return size_t(m_Resolution.x + (k == 0)) * size_t(m_Resolution.y + (k == 1)) *
               size_t(m_Resolution.z + (k == 2));
    }
    const Real* getConstGridDataXGPUPtr() const { return m_Data[0].data(); }
    const Real* getConstGridDataYGPUPtr() const { return m_Data[1].data(); }
    const Real* getConstGridDataZGPUPtr() const { return m_Data[2].data(); }
    Real* getGridDataXGPUPtr() { return m_Data[0].data(); }
    Real* getGridDataYGPUPtr() { return m_Data[1].data(); }
    Real* getGridDataZGPUPtr() { return m_Data[2].data(); }
    const Real* component(int k) const { return m_Data[k].data(); }
    Real* component(int k) { return m_Data[k].data(); }
    std::vector<Real> getHostData(int k) const { return m_Data[k].toHost(); }
    bool empty() const { return m_Data[0].size() == 0; }

private:
    DeviceArray m_Data[3];
};

inline void require_same_grid(const CGrid& a, const CGrid& b, const char* where) {
    if (a.getResolution() != b.getResolution())
        throw std::invalid_argument(std::string(where) + ": fields of different resolutions");
}

// ---------------------------------------------------------------------------
// CAdvectionSolver (MPhysics/AdvectionSolver.h:12-40): the virtual strategy
// slot of CGridFluidSolver::setAdvectionSolver.
class CAdvectionSolver {
public:
    virtual ~CAdvectionSolver() = default;
    virtual void advect(const CCellCenteredScalarField& vInputField, const CFaceCenteredVectorField& vVelocityField,
                        Real vDeltaT, CCellCenteredScalarField& voOutputField, EAdvectionAccuracy vEAdvectionAccuracy,
                        const CCellCenteredScalarField& vBoundarysSDF) = 0;
    virtual void advect(const CCellCenteredVectorField& vInputField, const CFaceCenteredVectorField& vVelocityField,
                        Real vDeltaT, CCellCenteredVectorField& voOutputField, EAdvectionAccuracy vEAdvectionAccuracy,
                        const CCellCenteredScalarField& vBoundarysSDF) = 0;
    virtual void advect(const CFaceCenteredVectorField& vInputField, const CFaceCenteredVectorField& vVelocityField,
                        Real vDeltaT, CFaceCenteredVectorField& voOutputField, EAdvectionAccuracy vEAdvectionAccuracy,
                        const CCellCenteredScalarField& vBoundarysSDF) = 0;
};

// CSemiLagrangian (MPhysics/SemiLagrangian.cpp:23-347): back-trace from the
// cell centres / face centres of the solver's grid, trilinear sample.
class CSemiLagrangian : public CAdvectionSolver {
public:
    CSemiLagrangian() = default;
    CSemiLagrangian(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0), const Vector3& h = Vector3(1, 1, 1)) {
        resizeSemiLagrangian(r, o, h);
    }
    void resizeSemiLagrangian(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                              const Vector3& h = Vector3(1, 1, 1)) {
        m_Grid = CCellCenteredScalarField(r, o, h);
        m_IsInit = true;
    }

    void advect(const CCellCenteredScalarField& in, const CFaceCenteredVectorField& vel, Real dt,
                CCellCenteredScalarField& out, EAdvectionAccuracy acc = EAdvectionAccuracy::RK2,
                const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) override {
        (void)sdf;
        if (!m_IsInit) return;   // an unsized solver does nothing, as the reference's m_IsInit test
        require_same_grid(in, vel, "CSemiLagrangian::advect");
        if (out.getResolution() != in.getResolution())
            out.resize(in.getResolution(), in.getOrigin(), in.getSpacing());
        const mgf_grid g = vel.toMgfGrid();
        check(mgf_op_advect_cc(ctx(), &g, in.getConstGridDataGPUPtr(), vel.getConstGridDataXGPUPtr(),
                               vel.getConstGridDataYGPUPtr(), vel.getConstGridDataZGPUPtr(), dt, int(acc),
                               out.getGridDataGPUPtr()),
              "mgf_op_advect_cc");
    }
    // CCV: each component sampled like a scalar at the same back-traced points
    // (sampleCellCenteredVectorFieldInvoker, MPhysics/FieldMathTool.cu:575-640)
    void advect(const CCellCenteredVectorField& in, const CFaceCenteredVectorField& vel, Real dt,
                CCellCenteredVectorField& out, EAdvectionAccuracy acc = EAdvectionAccuracy::RK2,
                const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) override {
        (void)sdf;
        if (!m_IsInit) return;
        require_same_grid(in, vel, "CSemiLagrangian::advect");
        if (out.getResolution() != in.getResolution())
            out.resize(in.getResolution(), in.getOrigin(), in.getSpacing());
        const mgf_grid g = vel.toMgfGrid();
        for (int k = 0; k < 3; ++k)
            check(mgf_op_advect_cc(ctx(), &g, in.component(k), vel.getConstGridDataXGPUPtr(),
                                   vel.getConstGridDataYGPUPtr(), vel.getConstGridDataZGPUPtr(), dt, int(acc),
                                   out.component(k)),
                  "mgf_op_advect_cc");
    }
    void advect(const CFaceCenteredVectorField& in, const CFaceCenteredVectorField& vel, Real dt,
                CFaceCenteredVectorField& out, EAdvectionAccuracy acc = EAdvectionAccuracy::RK2,
                const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) override {
        (void)sdf;
        if (!m_IsInit) return;
        require_same_grid(in, vel, "CSemiLagrangian::advect");
        if (out.getResolution() != in.getResolution())
            out.resize(in.getResolution(), in.getOrigin(), in.getSpacing());
        const mgf_grid g = vel.toMgfGrid();
        check(mgf_op_advect_fc(ctx(), &g, in.getConstGridDataXGPUPtr(), in.getConstGridDataYGPUPtr(),
                               in.getConstGridDataZGPUPtr(), vel.getConstGridDataXGPUPtr(),
                               vel.getConstGridDataYGPUPtr(), vel.getConstGridDataZGPUPtr(), dt, int(acc),
                               out.getGridDataXGPUPtr(), out.getGridDataYGPUPtr(), out.getGridDataZGPUPtr()),
              "mgf_op_advect_fc");
    }
    // CSemiLagrangian::backTrace: positions (CCV of x, y, z) moved back by dt.
    void backTrace(const CCellCenteredVectorField& inPos, const CFaceCenteredVectorField& vel, Real dt,
                   CCellCenteredVectorField& outPos, EAdvectionAccuracy acc = EAdvectionAccuracy::RK2) {
        if (&outPos != &inPos) outPos = inPos;
        const mgf_grid g = vel.toMgfGrid();
        check(mgf_op_backtrace(ctx(), &g, vel.getConstGridDataXGPUPtr(), vel.getConstGridDataYGPUPtr(),
                               vel.getConstGridDataZGPUPtr(), dt, int(acc), outPos.getGridDataXGPUPtr(),
                               outPos.getGridDataYGPUPtr(), outPos.getGridDataZGPUPtr(),
                               int64_t(outPos.getDataSize())),
              "mgf_op_backtrace");
    }

private:
    CCellCenteredScalarField m_Grid;   // the solver's grid (m_AdvectionInputPointPosFieldCC's)
    bool m_IsInit = false;
};

// CExternalForcesSolver (MPhysics/ExternalForcesSolver.cpp:18-59).
class CExternalForcesSolver {
public:
    CExternalForcesSolver() = default;
    CExternalForcesSolver(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0), const Vector3& h = Vector3(1, 1, 1)) {
        resizeExternalForcesSolver(r, o, h);
    }
    void resizeExternalForcesSolver(const Vector3i&, const Vector3& = Vector3(0, 0, 0),
                                    const Vector3& = Vector3(1, 1, 1)) {
        m_ExternalForces = Vector3(0, 0, 0);   // as the reference's resize resets them
    }
    void addExternalForces(Vector3 f) { m_ExternalForces += f; }
    void applyExternalForces(CFaceCenteredVectorField& vel, Real dt,
                             const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) {
        (void)sdf;
        const mgf_grid g = vel.toMgfGrid();
        const float gr[3] = {m_Gravity.x, m_Gravity.y, m_Gravity.z};
        const float fo[3] = {m_ExternalForces.x, m_ExternalForces.y, m_ExternalForces.z};
        check(mgf_op_apply_forces(ctx(), &g, vel.getGridDataXGPUPtr(), vel.getGridDataYGPUPtr(),
                                  vel.getGridDataZGPUPtr(), gr, fo, dt),
              "mgf_op_apply_forces");
    }
    const Vector3& getGravity() const { return m_Gravity; }
    const Vector3& getExternalForces() const { return m_ExternalForces; }

private:
    Vector3 m_Gravity = Vector3(0, Real(-9.8), 0);
    Vector3 m_ExternalForces = Vector3(0, 0, 0);
};

// ---------------------------------------------------------------------------
// EPGTransferAlgorithm (MPhysics/EnumType.h:19-26).
enum class EPGTransferAlgorithm {
    P2GSUM = MGF_PG_P2GSUM, G2PNEAREST = MGF_PG_G2PNEAREST, LINEAR = MGF_PG_LINEAR,
    QUADRATIC = MGF_PG_QUADRATIC, CUBIC = MGF_PG_CUBIC
};
using UInt = unsigned int;

// CEulerParticles (MPhysics/EulerParticles.h:5-64): a particle set in device
// memory the library owns.  Movable, not copyable (the reference's
// thrust-vector copies are replaced by explicit get/set through the host).
class CEulerParticles {
public:
    CEulerParticles() = default;
    explicit CEulerParticles(UInt n, const Real* pos = nullptr, const Real* vel = nullptr,
                             const Real* color = nullptr) {
        resize(n, pos, vel, color);
    }
    CEulerParticles(const CEulerParticles&) = delete;
    CEulerParticles& operator=(const CEulerParticles&) = delete;
    CEulerParticles(CEulerParticles&& o) noexcept : m_P(o.m_P) { o.m_P = nullptr; }
    CEulerParticles& operator=(CEulerParticles&& o) noexcept {
        std::swap(m_P, o.m_P);
        return *this;
    }
    ~CEulerParticles() { release(); }

    // CEulerParticles::resize (EulerParticles.cpp:20-67): host arrays or zeros.
    void resize(UInt n, const Real* pos = nullptr, const Real* vel = nullptr, const Real* color = nullptr) {
        release();
        check(mgf_particles_create(ctx(), int64_t(n), pos, vel, color, &m_P), "mgf_particles_create");
    }
    UInt getNumOfParticles() const {
        if (!m_P) return 0;
        int64_t n = 0;
        check(mgf_particles_count(m_P, &n), "mgf_particles_count");
        return UInt(n);
    }
    const Real* getConstParticlesPosGPUPtr() const { return array(MGF_PARTICLE_POS); }
    const Real* getConstParticlesVelGPUPtr() const { return array(MGF_PARTICLE_VEL); }
    const Real* getConstParticlesColorGPUPtr() const { return array(MGF_PARTICLE_COLOR); }
    Real* getParticlesPosGPUPtr() { return array(MGF_PARTICLE_POS); }
    Real* getParticlesVelGPUPtr() { return array(MGF_PARTICLE_VEL); }
    Real* getParticlesColorGPUPtr() { return array(MGF_PARTICLE_COLOR); }
    // getParticlesPos / Vel / Color as host copies.
    std::vector<Real> getParticlesPos() const { return fetch(MGF_PARTICLE_POS); }
    std::vector<Real> getParticlesVel() const { return fetch(MGF_PARTICLE_VEL); }
    std::vector<Real> getParticlesColor() const { return fetch(MGF_PARTICLE_COLOR); }
    void setParticlesPos(const Real* cpu) { put(MGF_PARTICLE_POS, cpu); }
    void setParticlesVel(const Real* cpu) { put(MGF_PARTICLE_VEL, cpu); }
    void setParticlesColor(const Real* cpu) { put(MGF_PARTICLE_COLOR, cpu); }

    // generateParticlesInFluid (EulerParticles.cpp:163-201): seeded jitter in
    // place of the reference's clock-seeded rand().
    void generateParticlesInFluid(const CCellCenteredScalarField& vFluidSDF, const CCellCenteredScalarField& vSolidSDF,
                                  UInt vNumOfPerGrid, uint64_t seed = 0) {
        require_same_grid(vFluidSDF, vSolidSDF, "CEulerParticles::generateParticlesInFluid");
        const mgf_grid g = vFluidSDF.toMgfGrid();
        mgf_particles p = nullptr;
        check(mgf_particles_generate(ctx(), &g, vFluidSDF.getConstGridDataGPUPtr(), vSolidSDF.getConstGridDataGPUPtr(),
                                     int32_t(vNumOfPerGrid), seed, &p),
              "mgf_particles_generate");
        release();
        m_P = p;
    }
    // statisticalFluidDensity (EulerParticles.cpp:203-206): particles per cell.
    void statisticalFluidDensity(CCellCenteredScalarField& vo) const {
        const mgf_grid g = vo.toMgfGrid();
        check(mgf_op_particle_markers(ctx(), &g, nullptr, getNumOfParticles(), getConstParticlesPosGPUPtr(),
                                      vo.getGridDataGPUPtr()),
              "mgf_op_particle_markers");
    }

    // The transfers of EulerParticles.cpp:208-261 (tranfer* invokers of
    // EulerParticlesTool.cu:443-868): particle values to grid fields,
    // normalised by the weight fields, and grid fields to particle values.
    void transferParticlesScalarValue2Field(CCellCenteredScalarField& vo, CCellCenteredScalarField& w,
                                            EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) const {
        p2g(vo, 0, array(MGF_PARTICLE_SCALAR), 1, 0, a, vo.getGridDataGPUPtr(), w, w.getGridDataGPUPtr());
    }
    void transferParticlesVectorValue2Field(CCellCenteredVectorField& vo, CCellCenteredVectorField& w,
                                            EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) const {
        for (int k = 0; k < 3; ++k)
            p2g(vo, 0, array(MGF_PARTICLE_VECTOR), 3, k, a, vo.component(k), w, w.component(k));
    }
    void transferParticlesVectorValue2Field(CFaceCenteredVectorField& vo, CFaceCenteredVectorField& w,
                                            EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) const {
        for (int k = 0; k < 3; ++k)
            p2g(vo, k + 1, array(MGF_PARTICLE_VECTOR), 3, k, a, vo.component(k), w, w.component(k));
    }
    void transferParticlesColor2Field(CCellCenteredVectorField& vo, CCellCenteredVectorField& w,
                                      EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) const {
        for (int k = 0; k < 3; ++k)
            p2g(vo, 0, array(MGF_PARTICLE_COLOR), 3, k, a, vo.component(k), w, w.component(k));
    }
    void transferParticlesVel2Field(CFaceCenteredVectorField& vo, CFaceCenteredVectorField& w,
                                    EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) const {
        for (int k = 0; k < 3; ++k)
            p2g(vo, k + 1, array(MGF_PARTICLE_VEL), 3, k, a, vo.component(k), w, w.component(k));
    }
    void transferScalarField2Particles(const CCellCenteredScalarField& f,
                                       EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) {
        g2p(f, 0, f.getConstGridDataGPUPtr(), a, MGF_PARTICLE_SCALAR, 1, 0);
    }
    void transferVectorField2Particles(const CCellCenteredVectorField& f,
                                       EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) {
        for (int k = 0; k < 3; ++k) g2p(f, 0, f.component(k), a, MGF_PARTICLE_VECTOR, 3, k);
    }
    void transferVectorField2Particles(const CFaceCenteredVectorField& f,
                                       EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) {
        for (int k = 0; k < 3; ++k) g2p(f, k + 1, f.component(k), a, MGF_PARTICLE_VECTOR, 3, k);
    }
    void transferVelField2Particles(const CFaceCenteredVectorField& f,
                                    EPGTransferAlgorithm a = EPGTransferAlgorithm::LINEAR) {
        for (int k = 0; k < 3; ++k) g2p(f, k + 1, f.component(k), a, MGF_PARTICLE_VEL, 3, k);
    }

    // advectParticlesInVelField (EulerParticles.cpp:231-325): a boundary SDF of
    // the velocity's resolution clamps the particles to the grid box.
    void advectParticlesInVelField(const CFaceCenteredVectorField& vel, Real dt, Real cfl,
                                   const CCellCenteredScalarField& boundary = CCellCenteredScalarField(),
                                   EAdvectionAccuracy acc = EAdvectionAccuracy::RK3) {
        const mgf_grid g = vel.toMgfGrid();
        const int fix = !boundary.empty() && boundary.getResolution() == vel.getResolution();
        check(mgf_particles_advect(handle(), &g, vel.getConstGridDataXGPUPtr(), vel.getConstGridDataYGPUPtr(),
                                   vel.getConstGridDataZGPUPtr(), dt, cfl, int32_t(acc), fix),
              "mgf_particles_advect");
    }

    mgf_particles handle() const {
        if (!m_P) throw std::logic_error("CEulerParticles: no particles (resize or generate first)");
        return m_P;
    }

private:
    Real* array(int which) const {
        float* p = nullptr;
        check(mgf_particles_data(handle(), which, &p), "mgf_particles_data");
        return p;
    }
    size_t count(int which) const { return size_t(getNumOfParticles()) * (which == MGF_PARTICLE_SCALAR ? 1 : 3); }
    std::vector<Real> fetch(int which) const {
        std::vector<Real> h(count(which));
        check(mgf_particles_download(handle(), which, h.data(), int64_t(h.size())), "mgf_particles_download");
        return h;
    }
    void put(int which, const Real* cpu) {
        check(mgf_particles_upload(handle(), which, cpu, int64_t(count(which))), "mgf_particles_upload");
    }
    void p2g(const CGrid& grid, int kind, const Real* values, int span, int off, EPGTransferAlgorithm a, Real* out,
             const CGrid& wgrid, Real* weight) const {
        require_same_grid(grid, wgrid, "CEulerParticles: field and weight field");
        const mgf_grid g = grid.toMgfGrid();
        check(mgf_op_particles_to_grid(ctx(), &g, kind, getNumOfParticles(), getConstParticlesPosGPUPtr(), values,
                                       span, off, int32_t(a), out, weight, 1),
              "mgf_op_particles_to_grid");
    }
    void g2p(const CGrid& grid, int kind, const Real* field, EPGTransferAlgorithm a, int which, int span, int off) {
        const mgf_grid g = grid.toMgfGrid();
        check(mgf_op_grid_to_particles(ctx(), &g, kind, getNumOfParticles(), getConstParticlesPosGPUPtr(), field,
                                       int32_t(a), array(which), span, off),
              "mgf_op_grid_to_particles");
    }
    void release() {
        if (m_P) mgf_particles_destroy(m_P);
        m_P = nullptr;
    }
    mgf_particles m_P = nullptr;
};

// The particle CAdvectionSolver plug-ins (CParticleInCell, ParticleInCell.cpp;
// CFluidImplicitParticle, FluidImplicitParticle.cpp; CMixPICAndFLIP,
// MixPICAndFLIP.cpp) share this base: particles generated in the fluid, the
// solver state of mgf_psolver, and the three advect overloads.
class CParticleAdvectionSolver : public CAdvectionSolver {
public:
    ~CParticleAdvectionSolver() override { release(); }
    CParticleAdvectionSolver(const CParticleAdvectionSolver&) = delete;
    CParticleAdvectionSolver& operator=(const CParticleAdvectionSolver&) = delete;

    void advect(const CCellCenteredScalarField& in, const CFaceCenteredVectorField& vel, Real dt,
                CCellCenteredScalarField& out, EAdvectionAccuracy acc = EAdvectionAccuracy::RK3,
                const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) override {
        (void)vel; (void)dt; (void)acc; (void)sdf;
        size_like(out, in);
        check(mgf_psolver_advect_scalar(solver(), in.getConstGridDataGPUPtr(), out.getGridDataGPUPtr()),
              "mgf_psolver_advect_scalar");
    }
    void advect(const CCellCenteredVectorField& in, const CFaceCenteredVectorField& vel, Real dt,
                CCellCenteredVectorField& out, EAdvectionAccuracy acc = EAdvectionAccuracy::RK3,
                const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) override {
        (void)vel; (void)dt; (void)acc; (void)sdf;
        size_like(out, in);
        check(mgf_psolver_advect_vector(solver(), in.component(0), in.component(1), in.component(2),
                                        out.component(0), out.component(1), out.component(2)),
              "mgf_psolver_advect_vector");
    }
    void advect(const CFaceCenteredVectorField& in, const CFaceCenteredVectorField& vel, Real dt,
                CFaceCenteredVectorField& out, EAdvectionAccuracy acc = EAdvectionAccuracy::RK3,
                const CCellCenteredScalarField& sdf = CCellCenteredScalarField()) override {
        size_like(out, in);
        const int fix = !sdf.empty() && sdf.getResolution() == vel.getResolution();
        if (&out == &in || &out == &vel) {   // the library's output may not alias its inputs
            CFaceCenteredVectorField tmp(in.getResolution(), in.getOrigin(), in.getSpacing());
            run_velocity(in, vel, dt, acc, fix, tmp);
            out = tmp;
            return;
        }
        run_velocity(in, vel, dt, acc, fix, out);
    }

    const CEulerParticles& getEulerParticles() const { return m_EulerParticles; }
    mgf_psolver solver() const {
        if (!m_S) throw std::logic_error("particle advection solver: not sized");
        return m_S;
    }
    // The generation seed (the reference seeds rand() with the clock).
    void setSeed(uint64_t seed) { m_Seed = seed; }

protected:
    explicit CParticleAdvectionSolver(int kind) : m_Kind(kind) {}
    void resize_solver(const CCellCenteredScalarField& fluid, const CCellCenteredScalarField& solid, UInt perGrid,
                       const Vector3i& r, const Vector3& o, const Vector3& h, Real cfl, Real mixing,
                       EPGTransferAlgorithm alg) {
        release();
        m_EulerParticles.generateParticlesInFluid(fluid, solid, perGrid, m_Seed);
        CCellCenteredScalarField g(r, o, h);
        const mgf_grid mg = g.toMgfGrid();
        check(mgf_psolver_create(ctx(), &mg, m_Kind, m_EulerParticles.handle(), cfl, int32_t(alg), mixing, &m_S),
              "mgf_psolver_create");
    }

private:
    template <class F> static void size_like(F& out, const F& in) {
        if (out.getResolution() != in.getResolution() || out.empty())
            out.resize(in.getResolution(), in.getOrigin(), in.getSpacing());
    }
    void run_velocity(const CFaceCenteredVectorField& in, const CFaceCenteredVectorField& vel, Real dt,
                      EAdvectionAccuracy acc, int fix, CFaceCenteredVectorField& out) {
        check(mgf_psolver_advect_velocity(solver(), in.component(0), in.component(1), in.component(2),
                                          vel.component(0), vel.component(1), vel.component(2), dt, int32_t(acc),
                                          fix, out.component(0), out.component(1), out.component(2)),
              "mgf_psolver_advect_velocity");
    }
    void release() {
        if (m_S) mgf_psolver_destroy(m_S);
        m_S = nullptr;
    }
    int m_Kind;
    uint64_t m_Seed = 0;
    CEulerParticles m_EulerParticles;
    mgf_psolver m_S = nullptr;
};

class CParticleInCell : public CParticleAdvectionSolver {
public:
    CParticleInCell() : CParticleAdvectionSolver(MGF_PSOLVER_PIC) {}
    CParticleInCell(const CCellCenteredScalarField& vFluidSDF, const CCellCenteredScalarField& vSolidSDF,
                    UInt vNumOfPerGrid, const Vector3i& vResolution, const Vector3& vOrigin = Vector3(0, 0, 0),
                    const Vector3& vSpacing = Vector3(1, 1, 1), Real vCFLNumber = 1.0,
                    EPGTransferAlgorithm vPGTransferAlg = EPGTransferAlgorithm::LINEAR)
        : CParticleAdvectionSolver(MGF_PSOLVER_PIC) {
        resizeParticleInCell(vFluidSDF, vSolidSDF, vNumOfPerGrid, vResolution, vOrigin, vSpacing, vCFLNumber,
                             vPGTransferAlg);
    }
    void resizeParticleInCell(const CCellCenteredScalarField& vFluidSDF, const CCellCenteredScalarField& vSolidSDF,
                              UInt vNumOfPerGrid, const Vector3i& vResolution,
                              const Vector3& vOrigin = Vector3(0, 0, 0), const Vector3& vSpacing = Vector3(1, 1, 1),
                              Real vCFLNumber = 1.0,
                              EPGTransferAlgorithm vPGTransferAlg = EPGTransferAlgorithm::LINEAR) {
        resize_solver(vFluidSDF, vSolidSDF, vNumOfPerGrid, vResolution, vOrigin, vSpacing, vCFLNumber, 0.0f,
                      vPGTransferAlg);
    }
};

class CFluidImplicitParticle : public CParticleAdvectionSolver {
public:
    CFluidImplicitParticle() : CParticleAdvectionSolver(MGF_PSOLVER_FLIP) {}
    CFluidImplicitParticle(const CCellCenteredScalarField& vFluidSDF, const CCellCenteredScalarField& vSolidSDF,
                           UInt vNumOfPerGrid, const Vector3i& vResolution, const Vector3& vOrigin = Vector3(0, 0, 0),
                           const Vector3& vSpacing = Vector3(1, 1, 1), Real vCFLNumber = 1.0,
                           EPGTransferAlgorithm vPGTransferAlg = EPGTransferAlgorithm::LINEAR)
        : CParticleAdvectionSolver(MGF_PSOLVER_FLIP) {
        resizeFluidImplicitParticle(vFluidSDF, vSolidSDF, vNumOfPerGrid, vResolution, vOrigin, vSpacing, vCFLNumber,
                                    vPGTransferAlg);
    }
    void resizeFluidImplicitParticle(const CCellCenteredScalarField& vFluidSDF,
                                     const CCellCenteredScalarField& vSolidSDF, UInt vNumOfPerGrid,
                                     const Vector3i& vResolution, const Vector3& vOrigin = Vector3(0, 0, 0),
                                     const Vector3& vSpacing = Vector3(1, 1, 1), Real vCFLNumber = 1.0,
                                     EPGTransferAlgorithm vPGTransferAlg = EPGTransferAlgorithm::LINEAR) {
        resize_solver(vFluidSDF, vSolidSDF, vNumOfPerGrid, vResolution, vOrigin, vSpacing, vCFLNumber, 0.0f,
                      vPGTransferAlg);
    }
};

class CMixPICAndFLIP : public CParticleAdvectionSolver {
public:
    CMixPICAndFLIP() : CParticleAdvectionSolver(MGF_PSOLVER_MIX) {}
    CMixPICAndFLIP(const CCellCenteredScalarField& vFluidSDF, const CCellCenteredScalarField& vSolidSDF,
                   UInt vNumOfPerGrid, const Vector3i& vResolution, const Vector3& vOrigin = Vector3(0, 0, 0),
                   const Vector3& vSpacing = Vector3(1, 1, 1), Real vCFLNumber = 1.0, Real vMixingCoefficient = 0.1f,
                   EPGTransferAlgorithm vPGTransferAlg = EPGTransferAlgorithm::LINEAR)
        : CParticleAdvectionSolver(MGF_PSOLVER_MIX) {
        resizeMixPICAndFLIP(vFluidSDF, vSolidSDF, vNumOfPerGrid, vResolution, vOrigin, vSpacing, vCFLNumber,
                            vMixingCoefficient, vPGTransferAlg);
    }
    void resizeMixPICAndFLIP(const CCellCenteredScalarField& vFluidSDF, const CCellCenteredScalarField& vSolidSDF,
                             UInt vNumOfPerGrid, const Vector3i& vResolution,
                             const Vector3& vOrigin = Vector3(0, 0, 0), const Vector3& vSpacing = Vector3(1, 1, 1),
                             Real vCFLNumber = 1.0, Real vMixingCoefficient = 0.1f,
                             EPGTransferAlgorithm vPGTransferAlg = EPGTransferAlgorithm::LINEAR) {
        resize_solver(vFluidSDF, vSolidSDF, vNumOfPerGrid, vResolution, vOrigin, vSpacing, vCFLNumber,
                      vMixingCoefficient, vPGTransferAlg);
    }
};

// Precision and stopping rule of a pressure solve.  The defaults are the
// reference's: plain FP32 CG, identity preconditioner, sqrt(||r||) < 1e-3,
// at most one iteration per cell (PressureSolver.cpp:80-82, CuMatrixFreePCG.h:27-28).
struct PressureOptions {
    int32_t precision = MGF_PRECISION_PLAIN;
    int32_t preconditioner = MGF_PRECOND_IDENTITY;
    int32_t tol_mode = MGF_TOL_ABS_SQRT_NORM;
    double tolerance = 1e-3;
    int64_t max_iterations = 0;
    bool warm_start = false;   // MIXED: start from the previous pressure (not the reference's x = 0)
};

inline void apply_options(mgf_scene_desc& d, const PressureOptions& p) {
    d.precision = p.precision;
    d.preconditioner = p.preconditioner;
    d.tol_mode = p.tol_mode;
    d.tolerance = p.tolerance;
    d.max_iterations = p.max_iterations;
    d.warm_start = p.warm_start ? 1 : 0;
}

// A scene of the library, owned.
class SceneHandle {
public:
    SceneHandle() = default;
    explicit SceneHandle(const mgf_scene_desc& d) { check(mgf_scene_create(ctx(), &d, &m_S), "mgf_scene_create"); }
    SceneHandle(const SceneHandle&) = delete;
    SceneHandle& operator=(const SceneHandle&) = delete;
    SceneHandle(SceneHandle&& o) noexcept : m_S(o.m_S) { o.m_S = nullptr; }
    SceneHandle& operator=(SceneHandle&& o) noexcept {
        std::swap(m_S, o.m_S);
        return *this;
    }
    ~SceneHandle() {
        if (m_S) mgf_scene_destroy(m_S);
    }
    mgf_scene get() const { return m_S; }
    explicit operator bool() const { return m_S != nullptr; }

private:
    mgf_scene m_S = nullptr;
};

inline void upload(mgf_scene s, int field, const Real* dev, size_t n) {
    check(mgf_scene_upload_device(s, field, dev, int64_t(n)), "mgf_scene_upload_device");
}
inline void download(mgf_scene s, int field, Real* dev, size_t n) {
    check(mgf_scene_download_device(s, field, dev, int64_t(n)), "mgf_scene_download_device");
}

inline mgf_scene_desc desc_for(Real dt, const Vector3i& r, const Vector3& o, const Vector3& h) {
    mgf_scene_desc d;
    const float origin[3] = {o.x, o.y, o.z}, spacing[3] = {h.x, h.y, h.z};
    check(mgf_scene_desc_defaults(&d, r.x, r.y, r.z, origin, spacing, dt), "mgf_scene_desc_defaults");
    return d;
}

// CPressureSolver (MPhysics/PressureSolver.cpp:85-164): markers, matrix,
// right-hand side, CG, pressure gradient, on the caller's fields.
class CPressureSolver {
public:
    CPressureSolver() = default;
    CPressureSolver(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0), const Vector3& h = Vector3(1, 1, 1)) {
        resizePressureSolver(r, o, h);
    }
    // The scene that holds the solve's buffers is made at the first solve.
    void resizePressureSolver(const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                              const Vector3& h = Vector3(1, 1, 1)) {
        m_Desc = desc_for(1.0f, r, o, h);
        apply_options(m_Desc, m_Options);
        m_Scene = SceneHandle();
        m_Sized = true;
    }
    void setOptions(const PressureOptions& p) {
        m_Options = p;
        apply_options(m_Desc, m_Options);
        if (m_Scene) check(mgf_scene_set_params(m_Scene.get(), &m_Desc), "mgf_scene_set_params");
    }
    const PressureOptions& getOptions() const { return m_Options; }

    void solvePressure(CFaceCenteredVectorField& vioFluidVelField, Real vDeltaT,
                       const CFaceCenteredVectorField& vSolidVelField, const CCellCenteredScalarField& vSolidSDFField,
                       const CCellCenteredScalarField& vFluidSDFField) {
        if (!m_Sized) throw std::logic_error("CPressureSolver: resizePressureSolver first");
        require_same_grid(vioFluidVelField, vFluidSDFField, "CPressureSolver::solvePressure");
        mgf_scene s = stage(vioFluidVelField, vDeltaT, vSolidVelField, vSolidSDFField);
        upload(s, MGF_FIELD_FLUID_DOMAIN, vFluidSDFField.getConstGridDataGPUPtr(), vFluidSDFField.getDataSize());
        const mgf_status st = mgf_scene_pressure_solve(s, &m_LastStats);
        // the reference ignores the CG result (PressureSolver.cpp:105) and projects anyway
        if (st != MGF_OK && st != MGF_ERR_NOT_CONVERGED) check(st, "mgf_scene_pressure_solve");
        for (int k = 0; k < 3; ++k)
            download(s, MGF_FIELD_VEL_X + k, vioFluidVelField.component(k), vioFluidVelField.faces(k));
    }

    // solvePressure with particle positions (PressureSolver.cpp:109-134): the
    // markers come from the particles.
    void solvePressure(CFaceCenteredVectorField& vioFluidVelField, Real vDeltaT,
                       const CFaceCenteredVectorField& vSolidVelField, const CCellCenteredScalarField& vSolidSDFField,
                       const CEulerParticles& vParticles) {
        if (!m_Sized) throw std::logic_error("CPressureSolver: resizePressureSolver first");
        require_same_grid(vioFluidVelField, vSolidSDFField, "CPressureSolver::solvePressure");
        mgf_scene s = stage(vioFluidVelField, vDeltaT, vSolidVelField, vSolidSDFField);
        const mgf_status st = mgf_scene_pressure_solve_particles(s, vParticles.handle(), &m_LastStats);
        if (st != MGF_OK && st != MGF_ERR_NOT_CONVERGED) check(st, "mgf_scene_pressure_solve_particles");
        for (int k = 0; k < 3; ++k)
            download(s, MGF_FIELD_VEL_X + k, vioFluidVelField.component(k), vioFluidVelField.faces(k));
    }

    // getMarkers / getConstVectorx / getConstVectorb (PressureSolver.h:24-30)
    // as cell-centred fields of the last solve.
    CCellCenteredScalarField getMarkers() const { return fetch(MGF_FIELD_MARKERS); }
    CCellCenteredScalarField getConstVectorx() const { return fetch(MGF_FIELD_PRESSURE); }
    CCellCenteredScalarField getConstVectorb() const { return fetch(MGF_FIELD_RHS); }
    const mgf_step_stats& getLastStats() const { return m_LastStats; }

private:
    // The solve's scene with the caller's velocity, solid velocity and solid marker.
    mgf_scene stage(const CFaceCenteredVectorField& vel, Real dt, const CFaceCenteredVectorField& svel,
                    const CCellCenteredScalarField& solid) {
        m_Desc.dt = dt;
        if (!m_Scene)
            m_Scene = SceneHandle(m_Desc);
        else
            check(mgf_scene_set_params(m_Scene.get(), &m_Desc), "mgf_scene_set_params");
        mgf_scene s = m_Scene.get();
        for (int k = 0; k < 3; ++k) {
            upload(s, MGF_FIELD_VEL_X + k, vel.component(k), vel.faces(k));
            upload(s, MGF_FIELD_SOLID_VEL_X + k, svel.component(k), svel.faces(k));
        }
        upload(s, MGF_FIELD_SOLID_SDF, solid.getConstGridDataGPUPtr(), solid.getDataSize());
        return s;
    }
    CCellCenteredScalarField fetch(int field) const {
        if (!m_Scene) throw std::logic_error("CPressureSolver: no solve yet");
        const Vector3i r(m_Desc.nx, m_Desc.ny, m_Desc.nz);
        CCellCenteredScalarField f(r, Vector3(m_Desc.origin[0], m_Desc.origin[1], m_Desc.origin[2]),
                                   Vector3(m_Desc.spacing[0], m_Desc.spacing[1], m_Desc.spacing[2]));
        download(m_Scene.get(), field, f.getGridDataGPUPtr(), f.getDataSize());
        return f;
    }
    mgf_scene_desc m_Desc{};
    PressureOptions m_Options;
    SceneHandle m_Scene;
    bool m_Sized = false;
    mgf_step_stats m_LastStats{};
};

using AdvectionSolverPtr = std::shared_ptr<CAdvectionSolver>;
using ExternalForcesSolverPtr = std::shared_ptr<CExternalForcesSolver>;
using PressureSolverPtr = std::shared_ptr<CPressureSolver>;

// CGridFluidSolver (MPhysics/GridFluidSolver.cpp:27-555): one scene of the
// library.  With a CSemiLagrangian advection solver the whole step runs in the
// library (mgf_step); any other CAdvectionSolver is called through its virtual
// advect on the scene's fields, as the reference's strategy slot works.
class CGridFluidSolver {
public:
    CGridFluidSolver() = default;
    CGridFluidSolver(Real vDeltaT, const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                     const Vector3& h = Vector3(1, 1, 1)) {
        resizeGridFluidSolver(vDeltaT, r, o, h);
    }

    void resizeGridFluidSolver(Real vDeltaT, const Vector3i& r, const Vector3& o = Vector3(0, 0, 0),
                               const Vector3& h = Vector3(1, 1, 1)) {
        m_DeltaT = vDeltaT;
        m_Resolution = r; m_Origin = o; m_Spacing = h;
        m_CurSimulationTime = 0;
        m_Desc = desc_for(vDeltaT, r, o, h);
        m_Scene = SceneHandle(m_Desc);
        m_HasColor = false;
        m_AdvectionSolver = std::make_shared<CSemiLagrangian>(r, o, h);
        m_ExternalForcesSolver = std::make_shared<CExternalForcesSolver>(r, o, h);
        m_PressureSolver = std::make_shared<CPressureSolver>(r, o, h);
    }

    // The configuration's options beyond the reference's (precision, preconditioner, stop rule).
    void setPressureOptions(const PressureOptions& p) { m_PressureSolver->setOptions(p); }
    void setAdvectionAccuracy(EAdvectionAccuracy a) { m_Accuracy = a; }

    void addSolidBoundary(const CCellCenteredScalarField& sdf) {
        const std::vector<Real> h = sdf.getHostData();
        const float v0[3] = {0, 0, 0};
        check(mgf_scene_add_solid(scene(), h.data(), int64_t(h.size()), v0), "mgf_scene_add_solid");
    }
    void setSolidVelField(const CFaceCenteredVectorField& v) {
        for (int k = 0; k < 3; ++k) upload(scene(), MGF_FIELD_SOLID_VEL_X + k, v.component(k), v.faces(k));
    }
    void addExternalForce(Vector3 f) { m_ExternalForcesSolver->addExternalForces(f); }
    void setColorField(const CCellCenteredVectorField& c) {
        for (int k = 0; k < 3; ++k) upload(scene(), MGF_FIELD_COLOR_X + k, c.component(k), c.getDataSize());
        m_HasColor = true;
    }
    void setFluidDomainField(const CCellCenteredScalarField& f) {
        upload(scene(), MGF_FIELD_FLUID_DOMAIN, f.getConstGridDataGPUPtr(), f.getDataSize());
    }
    void setVelocityField(const CFaceCenteredVectorField& v) {
        for (int k = 0; k < 3; ++k) upload(scene(), MGF_FIELD_VEL_X + k, v.component(k), v.faces(k));
    }
    void setAdvectionSolver(const AdvectionSolverPtr& p) { m_AdvectionSolver = p; }
    void setExternalForcesSolver(const ExternalForcesSolverPtr& p) { m_ExternalForcesSolver = p; }
    void setPressureSolver(const PressureSolverPtr& p) { m_PressureSolver = p; }

    const CCellCenteredScalarField& getSolidSDFField() const {
        refresh(m_SolidView, MGF_FIELD_SOLID_SDF);
        return m_SolidView;
    }
    Real getCurSimulationTime() const { return m_CurSimulationTime; }
    const CCellCenteredVectorField& getColorField() const {
        if (m_ColorView.empty() || m_ColorView.getResolution() != m_Resolution)
            m_ColorView.resize(m_Resolution, m_Origin, m_Spacing);
        for (int k = 0; k < 3; ++k)
            download(scene(), MGF_FIELD_COLOR_X + k, m_ColorView.component(k), m_ColorView.getDataSize());
        return m_ColorView;
    }
    const CCellCenteredScalarField& getFluidDomainField() const {
        refresh(m_DomainView, MGF_FIELD_FLUID_DOMAIN);
        return m_DomainView;
    }
    const CFaceCenteredVectorField& getVelocityField() const {
        if (m_VelView.empty() || m_VelView.getResolution() != m_Resolution)
            m_VelView.resize(m_Resolution, m_Origin, m_Spacing);
        for (int k = 0; k < 3; ++k)
            download(scene(), MGF_FIELD_VEL_X + k, m_VelView.component(k), m_VelView.faces(k));
        return m_VelView;
    }
    const AdvectionSolverPtr& getAdvectionSolver() const { return m_AdvectionSolver; }
    const PressureSolverPtr& getPressureSolver() const { return m_PressureSolver; }
    const mgf_step_stats& getLastStats() const { return m_LastStats; }

    // CGridFluidSolver::update (GridFluidSolver.cpp:336-370).
    // _computePressure's typeid dispatch (GridFluidSolver.cpp:434-525): a
    // particle advection solver makes the markers from its particles, and the
    // whole step then runs in the library (mgf_step_particles).
    void update() {
        prepare();
        if (fast_path()) {
            const mgf_status st = mgf_step(scene(), 1, &m_LastStats);
            if (st != MGF_OK && st != MGF_ERR_NOT_CONVERGED) check(st, "mgf_step");
            m_CurSimulationTime += m_DeltaT;
            return;
        }
        if (const CParticleAdvectionSolver* ps = particle_solver()) {
            const mgf_status st = mgf_step_particles(scene(), ps->solver(), 1, &m_LastStats);
            if (st != MGF_OK && st != MGF_ERR_NOT_CONVERGED) check(st, "mgf_step_particles");
            m_CurSimulationTime += m_DeltaT;
            return;
        }
        updateWithoutPressure();
        solvePressure();
    }
    // updateWithoutPressure / solvePressure (GridFluidSolver.cpp:341-357).
    void updateWithoutPressure() {
        prepare();
        check(mgf_scene_update_boundaries(scene()), "mgf_scene_update_boundaries");
        if (fast_path())
            check(mgf_scene_advect(scene()), "mgf_scene_advect");
        else
            advect_through_solver();
        check(mgf_scene_apply_forces(scene()), "mgf_scene_apply_forces");
    }
    void solvePressure() {
        const CParticleAdvectionSolver* ps = particle_solver();
        const mgf_status st = ps ? mgf_scene_pressure_solve_particles(scene(), ps->getEulerParticles().handle(),
                                                                      &m_LastStats)
                                 : mgf_scene_pressure_solve(scene(), &m_LastStats);
        if (st != MGF_OK && st != MGF_ERR_NOT_CONVERGED) check(st, "mgf_scene_pressure_solve");
        check(mgf_scene_extrapolate_velocity(scene(), &m_LastStats.extrapolation_layers),
              "mgf_scene_extrapolate_velocity");
        m_CurSimulationTime += m_DeltaT;
    }

    // generateCurFluidSDF / generateCurFluidDensity (GridFluidSolver.cpp:571-613)
    // fill the fields getFluidSDFField / getFluidDensityField return.
    void generateCurFluidSDF() { refresh(m_FluidSDFField, MGF_FIELD_FLUID_SDF); }
    void generateCurFluidDensity() { refresh(m_FluidDensityField, MGF_FIELD_FLUID_DENSITY); }
    const CCellCenteredScalarField& getFluidSDFField() const { return m_FluidSDFField; }
    const CCellCenteredScalarField& getFluidDensityField() const { return m_FluidDensityField; }

    void generateFluidDomainFromBBox(Vector3 mn, Vector3 mx) {
        const float a[3] = {mn.x, mn.y, mn.z}, b[3] = {mx.x, mx.y, mx.z};
        check(mgf_scene_fluid_box(scene(), a, b), "mgf_scene_fluid_box");
    }

    mgf_scene scene() const {
        if (!m_Scene) throw std::logic_error("CGridFluidSolver: resizeGridFluidSolver first");
        return m_Scene.get();
    }

private:
    bool fast_path() const { return dynamic_cast<CSemiLagrangian*>(m_AdvectionSolver.get()) != nullptr; }
    const CParticleAdvectionSolver* particle_solver() const {
        return dynamic_cast<const CParticleAdvectionSolver*>(m_AdvectionSolver.get());
    }

    // The scene's parameters follow the current solvers: forces, pressure options, accuracy.
    void prepare() {
        // _onBeginAdvanceTimeStep exits on a null solver (GridFluidSolver.cpp:374-388); here it throws
        if (!m_AdvectionSolver || !m_ExternalForcesSolver || !m_PressureSolver)
            throw std::logic_error("CGridFluidSolver: a solver is null");
        mgf_scene_desc d = m_Desc;
        const Vector3& g = m_ExternalForcesSolver->getGravity();
        const Vector3& f = m_ExternalForcesSolver->getExternalForces();
        d.gravity[0] = g.x; d.gravity[1] = g.y; d.gravity[2] = g.z;
        d.external_force[0] = f.x; d.external_force[1] = f.y; d.external_force[2] = f.z;
        d.rk_order = int32_t(m_Accuracy);
        apply_options(d, m_PressureSolver->getOptions());
        check(mgf_scene_set_params(scene(), &d), "mgf_scene_set_params");   // host-side only
        m_Desc = d;
    }

    // _computeAdvection (GridFluidSolver.cpp:399-417) through the virtual
    // strategy: velocity, colour, fluid domain, all with the old velocity.
    void advect_through_solver() {
        const CFaceCenteredVectorField vel = getVelocityField();
        const CCellCenteredScalarField& sdf = getSolidSDFField();
        CFaceCenteredVectorField nvel(m_Resolution, m_Origin, m_Spacing);
        m_AdvectionSolver->advect(vel, vel, m_DeltaT, nvel, m_Accuracy, sdf);
        if (m_HasColor) {
            const CCellCenteredVectorField col = getColorField();
            CCellCenteredVectorField ncol(m_Resolution, m_Origin, m_Spacing);
            m_AdvectionSolver->advect(col, vel, m_DeltaT, ncol, m_Accuracy, sdf);
            for (int k = 0; k < 3; ++k) upload(scene(), MGF_FIELD_COLOR_X + k, ncol.component(k), ncol.getDataSize());
        }
        const CCellCenteredScalarField dom = getFluidDomainField();
        CCellCenteredScalarField ndom(m_Resolution, m_Origin, m_Spacing);
        m_AdvectionSolver->advect(dom, vel, m_DeltaT, ndom, m_Accuracy, sdf);
        setVelocityField(nvel);
        setFluidDomainField(ndom);
    }

    void refresh(CCellCenteredScalarField& f, int field) const {
        if (f.empty() || f.getResolution() != m_Resolution) f.resize(m_Resolution, m_Origin, m_Spacing);
        download(scene(), field, f.getGridDataGPUPtr(), f.getDataSize());
    }

    Vector3i m_Resolution;
    Vector3 m_Origin, m_Spacing;
    Real m_DeltaT = 0;
    Real m_CurSimulationTime = 0;
    EAdvectionAccuracy m_Accuracy = EAdvectionAccuracy::RK3;   // _computeAdvection's RK3
    bool m_HasColor = false;
    mgf_scene_desc m_Desc{};
    SceneHandle m_Scene;
    AdvectionSolverPtr m_AdvectionSolver;
    ExternalForcesSolverPtr m_ExternalForcesSolver;
    PressureSolverPtr m_PressureSolver;
    mgf_step_stats m_LastStats{};
    CCellCenteredScalarField m_FluidSDFField, m_FluidDensityField;
    mutable CCellCenteredScalarField m_SolidView, m_DomainView;
    mutable CCellCenteredVectorField m_ColorView;
    mutable CFaceCenteredVectorField m_VelView;
};

}  // namespace mgf
