// mgf_internal.h — host-side state of the library and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <mutex>
#include <vector>


#include "../../include/mgf.h"
#include "mgf_device.h"

namespace mgf {

// Device scalars of the fused CG (one struct per scene, in device memory).
struct CgScalars {
    double rz;        // r . z of the current iteration
    double rz_new;
    double pap;       // p . A p
    double alpha;
    double beta;
    double rr;        // ||r||^2 (recurrence)
    double target2;   // stop when rr <= target2
    double sweep_r0;  // ||r||^2 at the start of the sweep
    int32_t done;     // 1 once rr <= target2 (or pap == 0)
    int32_t iters;    // iterations completed in this sweep
    int32_t pad0, pad1;
};

struct ReduceBuf {
    double* partials = nullptr;     // block partial sums (two per block at most)
    double* out2 = nullptr;         // two finished sums (refinement residual, ||b||^2)
};

// One multigrid level: extents and the per-axis operator scales.
struct MgDims {
    int nx, ny, nz;
    float sx, sy, sz;
};

// Device buffers of a coarse multigrid level (level 0 is the pressure grid).
struct MgLevelBuf {
    MgDims d{};
    uint8_t* markers = nullptr;
    uint16_t* codes = nullptr;
    float* b = nullptr;
    float* x = nullptr;
    float* t = nullptr;     // Jacobi ping-pong partner of x
    uint32_t* sched = nullptr; // fused-march schedule (this level as the fine one)
    size_t sched_n = 0;
    double march_cells = -1.0; // cells its schedule marches (measurement; < 0: unknown)
};

// Kernel probes (measurement).  While a probed scene's entry point runs, its
// sink is installed for the host thread and every kernel launch records a
// pair of HIP events at the kernel's own start and end
// (hipExtLaunchKernelGGL), with the kernel's name and the algorithmic bytes
// the launch site gave for it (probe_next_bytes; 0 when none was given).
struct ProbeSink {
    std::vector<hipEvent_t> ev;          // event pool, two per record
    std::vector<const char*> name;       // per record since the last collection
    std::vector<double> bytes;
    size_t used = 0;
    // collected records (durations read after a synchronisation)
    std::vector<std::string> done_name;
    std::vector<double> done_ms, done_bytes;
    // the next unused event pair; false when none can be created
    bool next_pair(hipEvent_t& a, hipEvent_t& b);
    void collect();                      // the stream has been synchronised
    void release();
};

struct Solid {
    float* sdf0 = nullptr;   // original SDF on the device
    float vel[3] = {0, 0, 0};
    float trans[3] = {0, 0, 0};
};

}  // namespace mgf

namespace mgf {
// Device buffers, pinned host buffers and streams of the lockstep batch
// plans, kept by the context and reused from call to call (a served loop
// builds plans every round): no allocation, free or stream creation after
// the first rounds, and no free that waits on the device.
struct CtxPool {
    struct Buf {
        void* p;
        size_t bytes;
        bool pinned;
        bool used;
    };
    std::mutex mu;
    std::vector<Buf> bufs;
    std::vector<hipStream_t> streams;   // free streams (default priority)
};
}  // namespace mgf

struct mgf_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 65536;
    bool lockstep = true;   // mgf_batch_step runs eligible scenes in lockstep
    mgf::CtxPool pool;
};

struct mgf_scene_s {
    mgf_context ctx = nullptr;
    hipStream_t stream = nullptr;  // the scene's own stream: scenes overlap on the device
    bool dry = false;              // sizing only: allocations are counted, not made
    mgf_scene_desc desc{};
    mgf::GridDev g{};
    size_t ncell = 0, nfx = 0, nfy = 0, nfz = 0, nvec = 0;

    // current velocity / fluid domain are u[cur] etc.; advection writes 1-cur.
    float* u[2] = {nullptr, nullptr};
    float* v[2] = {nullptr, nullptr};
    float* w[2] = {nullptr, nullptr};
    float* phi[2] = {nullptr, nullptr};
    float* color[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};  // from the first upload on
    int cur = 0;

    float* solid_total = nullptr;  // CBoundarys total marker: -k or UNKNOWN (0 with no solid)
    uint8_t* solid_mask = nullptr; // 1 where solid_total < 0
    float* us = nullptr;
    float* vs = nullptr;
    float* ws = nullptr;
    uint8_t* markers = nullptr;    // 0 air / 1 fluid / 2 solid

    // pressure system
    float* b = nullptr;
    float* x = nullptr;            // FP32 pressure (plain) / inner correction (mixed)
    double* x64 = nullptr;         // FP64 pressure (mixed)
    float* r = nullptr;
    float* p0 = nullptr;
    float* p1 = nullptr;
    float* ap = nullptr;           // plain: A p
    mgf::CgScalars* scal = nullptr;
    mgf::ReduceBuf red;
    double* host_scratch = nullptr; // pinned host mirror of small reductions
    uint16_t* codes = nullptr;     // per-cell stencil codes of the pressure matrix
    // preconditioner (allocated on first use)
    float* z = nullptr;            // M^-1 r
    float* zt = nullptr;           // Jacobi ping-pong partner of z
    uint32_t* sched0 = nullptr;    // fused-march schedule of level 0
    size_t sched0_n = 0;
    uint8_t* wet_flat = nullptr;   // blocks of the flat CG kernels with fluid
    // with probes on: cells in blocks with fluid (flat kernels, level-0
    // march) for the bytes a probed launch moved; < 0 = every cell
    double wet_flat_cells = -1.0, wet0_cells = -1.0;
    std::vector<mgf::MgLevelBuf> mg;   // coarse levels 1..L-1
    int mg_nx = -1, mg_ny = -1, mg_nz = -1;
    int mg_l0 = 0;                 // levels [mg_l0, L) run in one work-group
    bool x_is_64 = false;          // which pressure the last solve produced
    bool x64_valid = false;        // x64 holds a finished mixed solve (a warm start's guess)

    // extrapolation scratch
    int32_t* layer = nullptr;      // nfx+nfy+nfz words: the faces' uint16 distances, or the SDF's scratch
    uint32_t* list = nullptr;      // faces sorted by distance (nfx+nfy+nfz)
    uint8_t* lmask = nullptr;      // previous-layer neighbour mask of each list entry
    uint8_t* fmask = nullptr;      // the same mask per face (read for neighbours of the next layer)
    uint32_t* counts = nullptr;    // counting-sort scratch [ext_counts_size(K)]
    uint32_t* host_offs = nullptr; // pinned mirror of the offsets and the max distance

    std::vector<mgf::Solid> solids;
    bool boundary_dirty = false;   // solids changed since the last refresh
    bool solids_unbuilt = false;   // solids added since the marker was last built
    bool boundary_user = false;    // SOLID_SDF / SOLID_VEL uploaded directly
    bool timing = false;
    // kernel probes: every launch of the scene's entry points while enabled
    bool probes = false;
    mgf::ProbeSink sink;
    bool skip_dry = true;          // CG / fused multigrid blocks without fluid return at once
    bool div_check = false;        // divergence check after each projection
    double* div_dev = nullptr;     // [0] largest ratio of the call, [1..2] the last sums

    int64_t device_bytes = 0;
    std::vector<void*> allocations;
};

// A particle set: interleaved xyz per particle; scratch = the advection
// temporary (the sampled velocity, 3n floats).
struct mgf_particles_s {
    mgf_context ctx = nullptr;
    int64_t n = 0;
    float* pos = nullptr;
    float* vel = nullptr;
    float* color = nullptr;
    float* scalar = nullptr;   // n
    float* vector = nullptr;   // 3n
    float* scratch = nullptr;  // 3n: the velocities sampled at the particles
    uint32_t* vmax = nullptr;  // max |v| bits
    uint64_t pos_version = 0;  // bumped whenever the positions change
    std::vector<void*> allocations;
};

// A particle advection solver on a borrowed particle set.
struct mgf_psolver_s {
    mgf_context ctx = nullptr;
    mgf_particles p = nullptr;
    mgf_grid grid{};
    int kind = 0, alg = 0;
    float cfl = 1.0f, mixing = 0.1f;
    float* delta[3] = {nullptr, nullptr, nullptr};   // previous output velocity / grid change (FLIP, MIX)
    float* weight[3] = {nullptr, nullptr, nullptr};  // weight fields (face extents; cells fit in them)
    // the particles binned by support corner for the splat (valid for bins_version)
    uint32_t *bin_counts = nullptr, *bin_tiles = nullptr, *bin_start = nullptr, *bin_cursor = nullptr;
    uint32_t* order = nullptr;
    float* acc = nullptr;                            // the splat's per-bin slots (LINEAR)
    float* pos_s = nullptr;                          // positions in bin order (3n)
    float* vel_s = nullptr;                          // velocities in bin order (3n)
    uint64_t bins_version = ~uint64_t(0);
    bool scalar_seeded = false, vector_seeded = false, vel_seeded = false;
    std::vector<void*> allocations;
};

namespace mgf {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 4096;   // fused reductions use at most this many blocks

// Error helpers.
void set_error(const std::string& msg);

ProbeSink*& probe_sink();
// Algorithmic bytes of the next kernel launched by this host thread.
void probe_next_bytes(double bytes);
double probe_take_bytes();

// Device bytes a scene with this descriptor holds once it has stepped (the
// preconditioner's levels included, solids excluded); nothing is allocated.
mgf_status scene_bytes_for(mgf_context ctx, const mgf_scene_desc* desc, int64_t* out);
// Makes the allocations a scene's first step would make, so that stepping
// allocates nothing.
mgf_status scene_prepare(mgf_scene s);
// Whether a scene can be stepped in lockstep with others (mixed precision,
// multigrid preconditioner, cold start).
bool lockstep_eligible(mgf_scene s);
// The context's pool: a buffer of at least `bytes` (device or pinned host),
// back with pool_put; a stream, back with pool_stream_put; pool_release
// frees everything (at mgf_shutdown).
void* pool_get(mgf_context ctx, size_t bytes, bool pinned);
void pool_put(mgf_context ctx, void* p);
hipStream_t pool_stream(mgf_context ctx);
void pool_stream_put(mgf_context ctx, hipStream_t s);
void pool_release(mgf_context ctx);
// nsteps steps of every scene in sc (all eligible, distinct, of ctx): the
// pressure solves as multi-scene launches on one stream, the other stages on
// the scenes' streams from up to `workers` host threads.  stats, status and
// errs receive one entry per scene; step_ms (n * nsteps, scene-major, may be
// null) the wall time of each batch step.
mgf_status batch_lockstep(mgf_context ctx, const std::vector<mgf_scene>& sc, int32_t nsteps, int32_t workers,
                          std::vector<mgf_step_stats>& stats, std::vector<mgf_status>& status,
                          std::vector<std::string>& errs, double* step_ms);
// Extrapolation buffers (face count F, K = ext_max_distance, below 65535):
// dist F, list F, lmask F, fmask F, counts ext_counts_size(K); host_offs is
// host memory of K + 3 words the layer offsets are read into.
struct ExtBufs {
    uint16_t* dist;
    uint32_t* list;
    uint8_t* lmask;
    uint8_t* fmask;
    uint32_t* counts;
    uint32_t* host_offs;
};
// extrapolatingDataInvoker(FCV) on (u, v, w) in place on stream st; *used
// receives the largest non-empty layer (< layers).  device_only: no host
// synchronisation (every layer in one work-group, the largest layer copied
// to host_offs[0] when the stream gets there; *used = -1 when kernels were
// enqueued).
mgf_status extrapolate_run(hipStream_t st, const GridDev& g, float* u, float* v, float* w, int layers,
                           const ExtBufs& B, int* used, bool device_only = false);
#define MGF_HIP(expr)                                                           \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::mgf::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return MGF_ERR_DEVICE;                                              \
        }                                                                       \
    } while (0)

inline GridDev make_grid(const mgf_grid* g) {
    GridDev d;
    d.nx = g->nx; d.ny = g->ny; d.nz = g->nz;
    d.ox = g->origin[0]; d.oy = g->origin[1]; d.oz = g->origin[2];
    d.hx = g->spacing[0]; d.hy = g->spacing[1]; d.hz = g->spacing[2];
    set_recip(d);
    return d;
}

RkCoef rk_coef(float dt);

// ---- launchers (mgf_kernels.hip) ------------------------------------------
hipError_t launch_backtrace(hipStream_t s, const GridDev& g, Vel V, float dt, int order,
                            float* px, float* py, float* pz, int64_t n);
hipError_t launch_sample_cc(hipStream_t s, const GridDev& g, const float* f, const float* px,
                            const float* py, const float* pz, float* out, int64_t n);
// Advects the three face components and applies the uniform increments
// inc0 (gravity*dt) and inc1 (force*dt) when enabled.
hipError_t launch_advect_fc(hipStream_t s, const GridDev& g, Vel in, Vel V, float dt, int order,
                            float* ou, float* ov, float* ow, int add0, const float inc0[3],
                            int add1, const float inc1[3]);
hipError_t launch_advect_cc(hipStream_t s, const GridDev& g, const float* in, Vel V, float dt,
                            int order, float* out);
// The fluid domain and the three colour components sampled at one back-traced
// point per cell.
hipError_t launch_advect_cc4(hipStream_t s, const GridDev& g, const float* phi, const float* const c[3], Vel V,
                             float dt, int order, float* ophi, float* const oc[3]);
hipError_t launch_add_uniform(hipStream_t s, float* a, size_t n, float inc);
hipError_t launch_markers(hipStream_t s, const GridDev& g, const uint8_t* solid_mask,
                          const float* fluid, uint8_t* markers);
hipError_t launch_rhs(hipStream_t s, const GridDev& g, Vel U, Vel S, const uint8_t* markers,
                      float* b);
hipError_t launch_matrix_a(hipStream_t s, const GridDev& g, const float scale[3],
                           const uint8_t* markers, float* A);
hipError_t launch_divergence(hipStream_t s, const GridDev& g, Vel U, float* div);
hipError_t launch_matvec(hipStream_t s, const GridDev& g, const float scale[3],
                         const uint8_t* markers, const float* x, float* y);
// Plain BLAS-1 (FP64 partial sums; result written to partials, reduced on host)
hipError_t launch_dot_partials(hipStream_t s, const float* a, const float* b, size_t n,
                               double* partials, int blocks);
hipError_t launch_axpy(hipStream_t s, float alpha, const float* x, float* y, size_t n);
hipError_t launch_scal_add(hipStream_t s, float beta, float* p, const float* z, size_t n);
hipError_t launch_fill(hipStream_t s, float* a, size_t n, float v);
hipError_t launch_fill64(hipStream_t s, double* a, size_t n, double v);
// FP64 ||b - A x|| and ||b|| partials for an FP32 pressure.
hipError_t launch_true_residual32(hipStream_t s, const GridDev& g, const float scale[3],
                                  const uint8_t* markers, const float* x, const float* b,
                                  double* partials, int blocks);
hipError_t launch_project(hipStream_t s, const GridDev& g, const float sg[3],
                          const uint8_t* markers, const float* p32, const double* p64,
                          float* u, float* v, float* w, const float* us, const float* vs,
                          const float* ws);
// Divergence check of the projected velocity: out[0] = max(out[0], ||D u|| /
// ||b||) over the fluid cells (FP64; out[1], out[2] receive the two sums);
// partials holds 2 * blocks doubles.
hipError_t launch_div_check(hipStream_t s, const GridDev& g, const uint8_t* markers, const float* u,
                            const float* v, const float* w, const float* b, double* partials, int blocks,
                            double* out);
// ---- column-marching CG and multigrid (mgf_kernels_mg.hip) -----------------
hipError_t launch_codes(hipStream_t s, int nx, int ny, int nz, const uint8_t* markers, uint16_t* codes);
// Blocks of the column kernels on grid g (the partial-sum buffer must hold
// four doubles per block).
uint32_t col_blocks(const GridDev& g);
// wet: per block of the flat kernels, whether it holds a fluid cell
// (launch_flat_wet, once per set of codes); null: no block skips.
hipError_t launch_flat_wet(hipStream_t s, const GridDev& g, const uint16_t* codes, uint8_t* wet);
// Cells in the blocks a host copy of the wet bytes marks (measurement).
double flat_wet_cells(const GridDev& g, const uint8_t* wet_host);
// K_I: p = z + beta p, Ap = A z + beta Ap in place (p = z, Ap = A z on the
// first iteration), p . Ap, alpha.
hipError_t launch_col_p_ap(hipStream_t s, const GridDev& g, const float scale[3], const uint16_t* codes,
                           const float* zv, float* p, float* ap, CgScalars* sc, ReduceBuf red, const uint8_t* wet);
// K_II: x += alpha p, r -= alpha Ap, r . r; identity != 0: beta = rr / rz, rz = rr.
hipError_t launch_col_update(hipStream_t s, const GridDev& g, const float* p, const float* ap, float* x, float* r,
                             CgScalars* sc, ReduceBuf red, int identity, const uint8_t* wet);
// Start of an inner sweep: r = rsrc, x = 0, the scalars from r.r (target =
// max(rel_inner^2 r.r, abs_target2)).
hipError_t launch_cg_start(hipStream_t s, const GridDev& g, const float* rsrc, float* r, float* x, CgScalars* sc,
                           ReduceBuf red, double rel_inner, double abs_target2);
// Refinement: mode 0: x64 = 0; mode 1: x64 = xcorr; mode 2: x64 += xcorr;
// mode 3: x64 kept on fluid cells, zeroed elsewhere (a warm start); then
// (modes 1-3) r = b - A x64 in FP64, stored FP32 when r != null.  out2 (device)
// receives ||b - A x64||^2 (0 in mode 0) and ||b||^2.
hipError_t launch_refine_true(hipStream_t s, const GridDev& g, const float scale[3], const uint16_t* codes,
                              const float* xcorr, double* x64, const float* b, float* r, int mode, ReduceBuf red,
                              double* out2);
hipError_t launch_mg_coarsen(hipStream_t s, const MgDims& f, const MgDims& c, const uint8_t* mf, uint8_t* mc,
                             int fluid_first);
hipError_t launch_mg_smooth(hipStream_t s, const MgDims& l, const uint16_t* code, const float* b, const float* xi,
                            float* xo, float omega, CgScalars* sc, ReduceBuf red, int first, int with_rz);
// The first two sweeps from x = 0 fused (x1 = omega D^-1 b is pointwise).
hipError_t launch_mg_smooth2_first(hipStream_t s, const MgDims& l, const uint16_t* code, const float* b, float* xo,
                                   float omega, CgScalars* sc, ReduceBuf red, int with_rz);
// A sweep on xi + (injection of the coarse xc) without storing the sum.
hipError_t launch_mg_smooth_prolong(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code,
                                    const float* b, const float* xi, const float* xc, float* xo, float omega,
                                    CgScalars* sc, ReduceBuf red, int with_rz);
// The V-cycle of levels[0..nlev) in one work-group with the iterates in LDS:
// b0 = right-hand side of levels[0] (global), x0 = its result (global).
// Fused descent (the first three sweeps from zero and the restriction of the
// residual) and ascent (prolongation and three sweeps; with_rz: r . z) of a
// level whose every axis is halved (mg_march_ok).  mg_march_grid = blocks of
// one launch; mg_march_parts = the partials of r . z the ascent writes.
bool mg_march_ok(const MgDims& f);
uint32_t mg_march_grid(const MgDims& f, const MgDims& c);
uint32_t mg_march_parts(const MgDims& f, const MgDims& c);
// The schedule of a fused level (32-bit words, mg_march_sched_words): made by
// launch_mg_march_wet and then launch_mg_march_plan once per set of codes,
// for `slots` resident blocks (two per CU).  It picks the planes per block
// (the record, for many scenes at once: the static chunk) and puts the
// blocks whose tile holds fluid first, evenly over the XCDs;
// with skip, the others store zeros and skip the march.  Null: the static
// split, every block marched.  Host copy: the marched cells (measurement).
size_t mg_march_sched_words(const MgDims& f, const MgDims& c);
hipError_t launch_mg_march_wet(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code,
                               uint32_t* sched);
hipError_t launch_mg_march_plan(hipStream_t s, const MgDims& f, const MgDims& c, uint32_t* sched, uint32_t slots);
double mg_march_sched_cells(const MgDims& f, const MgDims& c, const uint32_t* sched_host);
hipError_t launch_mg_descend(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                             float* x, const uint16_t* codec, float* bc, float omega, const CgScalars* sc,
                             const uint32_t* sched, int skip);
hipError_t launch_mg_ascend(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                            const float* xi, const float* xc, float* xo, float omega, CgScalars* sc, ReduceBuf red,
                            int with_rz, const uint32_t* sched, int skip);
size_t mg_coarse_lds_bytes(const MgDims* levels, int nlev);
hipError_t launch_mg_coarse_cycle(hipStream_t s, const MgDims* levels, const uint16_t* const* codes, int nlev,
                                  const float* b0, float* x0, int nu, int last_sweeps, float omega,
                                  const CgScalars* sc);
hipError_t launch_mg_restrict(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* codef,
                              const float* bf, const float* xf, const uint16_t* codec, float* bc,
                              const CgScalars* sc);
// ---- multi-scene launches (mgf_kernels_mg.hip) -----------------------------
// One scene's part of a multi-scene launch (k_multi): the kernel, its
// argument struct as bytes, and the block count and dynamic LDS of the
// scene's own launch.  The rec_* functions append the records of the launches
// the launch_* function of the same name makes, in its order.
enum MultiKind : int32_t {
    kMkStart, kMkFinish, kMkPap, kMkUpdate, kMkRefine1, kMkRefine2, kMkAcc64, kMkSum2,
    kMkSmoothTT, kMkSmoothTF, kMkSmoothFT, kMkSmoothFF, kMkSmooth2T, kMkSmooth2F, kMkProlongT, kMkProlongF,
    kMkRestrict, kMkRestrict2, kMkRestrictTile, kMkCoarse, kMkMarchDesc, kMkMarchAsc, kMkMarchAscRz, kMkGather,
    kMkRefine0, kMkCodes, kMkFlatWet, kMkCoarsen, kMkMarchWet, kMkMarchPlan,
    kMkAdvectFc, kMkAdvectCc, kMkAdvectCc4, kMkMarkers, kMkRhs, kMkProject, kMkCopyWord,
    kMkExtRows, kMkExtRowsReg4, kMkExtRowsReg8, kMkExtRowsReg16, kMkExtLines1, kMkExtLines2, kMkExtZero, kMkExtCount,
    kMkExtScan, kMkExtScatter, kMkExtTail, kMkCopyOffs, kMkExtRowsReg5,
    kMkExtLinesSeg,                          // 8 kinds: (axis - 1) * 4 + segment size class
    kMkExtLinesSegLast = kMkExtLinesSeg + 7,
    kMkExtRowsShort1, kMkExtRowsShort2,
    kMkExtLinesShort,                        // 4 kinds: (axis - 1) * 2 + segment size class, 4 waves a group
    kMkExtLinesShortLast = kMkExtLinesShort + 3,
    kMkCount
};
constexpr size_t kMRecBytes = 512;
// A multi-scene block map entry: (scene << kMapShift) | block of the scene's launch.
constexpr uint32_t kMapShift = 20;
constexpr uint32_t kMapLocal = (1u << kMapShift) - 1u;
struct MRec {
    int32_t kind = -1;
    uint32_t blocks = 0, shmem = 0, bytes = 0;
    alignas(16) unsigned char args[kMRecBytes];
};
void rec_cg_start(std::vector<MRec>& out, const GridDev& g, const float* rsrc, float* r, float* x, CgScalars* sc,
                  ReduceBuf red, double rel_inner, double abs_target2);
void rec_col_p_ap(std::vector<MRec>& out, const GridDev& g, const float scale[3], const uint16_t* codes,
                  const float* zv, float* p, float* ap, CgScalars* sc, ReduceBuf red, const uint8_t* wet);
void rec_col_update(std::vector<MRec>& out, const GridDev& g, const float* p, const float* ap, float* x, float* r,
                    CgScalars* sc, ReduceBuf red, int identity, const uint8_t* wet);
void rec_refine_true(std::vector<MRec>& out, const GridDev& g, const float scale[3], const uint16_t* codes,
                     const float* xcorr, double* x64, const float* b, float* r, int mode, ReduceBuf red,
                     double* out2);
void rec_mg_smooth(std::vector<MRec>& out, const MgDims& l, const uint16_t* code, const float* b, const float* xi,
                   float* xo, float omega, CgScalars* sc, ReduceBuf red, int first, int with_rz);
void rec_mg_smooth2_first(std::vector<MRec>& out, const MgDims& l, const uint16_t* code, const float* b, float* xo,
                          float omega, CgScalars* sc, ReduceBuf red, int with_rz);
void rec_mg_smooth_prolong(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code,
                           const float* b, const float* xi, const float* xc, float* xo, float omega, CgScalars* sc,
                           ReduceBuf red, int with_rz);
void rec_mg_restrict(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* codef, const float* bf,
                     const float* xf, const uint16_t* codec, float* bc, const CgScalars* sc);
bool rec_mg_coarse_cycle(std::vector<MRec>& out, const MgDims* levels, const uint16_t* const* codes, int nlev,
                         const float* b0, float* x0, int nu, int last_sweeps, float omega, const CgScalars* sc);
void rec_mg_descend(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                    float* x, const uint16_t* codec, float* bc, float omega, const CgScalars* sc,
                    const uint32_t* sched, int skip);
void rec_mg_ascend(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                   const float* xi, const float* xc, float* xo, float omega, CgScalars* sc, ReduceBuf red,
                   int with_rz, const uint32_t* sched, int skip);
// The finishing record's size, and its abs_target2 set in place.
size_t finish_arg_bytes();
void finish_set_target(unsigned char* args, double abs_target2);
void rec_gather(std::vector<MRec>& out, const CgScalars* sc, const double* out2, CgScalars* dst_sc, double* dst2);
void rec_codes(std::vector<MRec>& out, int nx, int ny, int nz, const uint8_t* markers, uint16_t* codes);
void rec_flat_wet(std::vector<MRec>& out, const GridDev& g, const uint16_t* codes, uint8_t* wet);
void rec_mg_coarsen(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint8_t* mf, uint8_t* mc,
                    int fluid_first);
void rec_mg_march_wet(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code,
                      uint32_t* sched);
void rec_mg_march_plan(std::vector<MRec>& out, const MgDims& f, const MgDims& c, uint32_t* sched);
void rec_advect_fc(std::vector<MRec>& out, const GridDev& g, Vel in, Vel V, float dt, int order, float* ou, float* ov,
                   float* ow, int add0, const float inc0[3], int add1, const float inc1[3]);
void rec_advect_cc(std::vector<MRec>& out, const GridDev& g, const float* in, Vel V, float dt, int order, float* o);
void rec_advect_cc4(std::vector<MRec>& out, const GridDev& g, const float* phi, const float* const c[3], Vel V,
                    float dt, int order, float* ophi, float* const oc[3]);
void rec_markers(std::vector<MRec>& out, const GridDev& g, const uint8_t* solid_mask, const float* fluid,
                 uint8_t* markers);
void rec_rhs(std::vector<MRec>& out, const GridDev& g, Vel U, Vel S, const uint8_t* markers, float* b);
void rec_project(std::vector<MRec>& out, const GridDev& g, const float sg[3], const uint8_t* markers, const float* p32,
                 const double* p64, float* u, float* v, float* w, const float* us, const float* vs, const float* ws);
void rec_copy_word(std::vector<MRec>& out, const uint32_t* src, uint32_t* dst);
// The extrapolation with no host synchronisation (distance transform, sort,
// every layer in one work-group; *maxd = the device address of the largest
// layer); K = ext_max_distance(g) < 65535, 1 <= cap <= K.
void rec_ext_device(std::vector<MRec>& out, const GridDev& g, float* u, float* v, float* w, uint16_t* dist,
                    int32_t cap, int32_t K, uint32_t* counts, uint32_t* list, uint8_t* lmask, uint8_t* fmask,
                    const uint32_t** maxd, bool with_layers = true);
// The sort's offsets 0..cap+1 and largest layer packed into dst (cap + 3 words).
void rec_copy_offs(std::vector<MRec>& out, uint32_t* counts, int32_t K, int32_t cap, uint32_t* dst);
// Layer L of many scenes in one launch (k_multi_layer): a table of
// layer_arg_bytes() per scene (layer_args), map entries (scene, block) with
// layer_entries_per_block() entries a block.
uint32_t layer_entries_per_block();
size_t layer_arg_bytes();
void layer_args(unsigned char* dst, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                const uint8_t* lmask, uint32_t* counts, int32_t K);
hipError_t launch_multi_layer(hipStream_t s, const void* tab, const uint32_t* map, uint32_t nblocks, int32_t L);
// The one-work-group layers first..last as a kMkExtTail record's arguments.
size_t tail_arg_bytes();
void tail_args(unsigned char* dst, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
               const uint8_t* lmask, uint32_t* counts, int32_t K, int32_t first, int32_t last);
hipError_t launch_multi_k(hipStream_t s, int32_t kind, const void* tab, const uint32_t* map, uint32_t nblocks,
                          uint32_t shmem, const uint8_t* active);
hipError_t launch_multi_ext(hipStream_t s, int32_t kind, const void* tab, const uint32_t* map, uint32_t nblocks,
                            uint32_t shmem, const uint8_t* active);
// One multi-scene launch: tab = the participating scenes' argument structs
// (one kind), map = per block (scene << 20) | block of the scene's launch;
// scenes whose active byte is 0 return at once (active may be null).
hipError_t launch_multi(hipStream_t s, int32_t kind, const void* tab, const uint32_t* map, uint32_t nblocks,
                        uint32_t shmem, const uint8_t* active);

// ---- extrapolation (mgf_kernels_ext.hip) ------------------------------------
// Largest L1 distance between two faces of one component grid (+ margin).
int32_t ext_max_distance(const GridDev& g);
// dist = L1 distance of every face to the nearest known face (65535 where
// a component has none; nx + ny + nz + 3 must be below 65535); faces indexed
// u, then v, then w.
hipError_t launch_ext_distance(hipStream_t s, const GridDev& g, const float* u, const float* v, const float* w,
                               uint16_t* dist);
// Counting sort of the faces with 1 <= dist <= cap (cap <= K) into list, by
// distance; lmask[k] = the mask of list[k]'s neighbours one layer below.  counts: ext_counts_size(K) entries of scratch; *offs receives
// the device address of offs[0..cap+1] (offs[L] = first slot of distance L,
// offs[cap+1] = total) and *maxd that of the largest distance present.
size_t ext_counts_size(int32_t K);
hipError_t launch_ext_sort(hipStream_t s, const GridDev& g, const uint16_t* dist, int32_t cap, int32_t K,
                           uint32_t* counts, uint32_t* list, uint8_t* lmask, uint8_t* fmask, const uint32_t** offs,
                           const uint32_t** maxd);
// One layer: list[begin, begin+count) take the mean of their neighbours in
// the previous layer (lmask).
hipError_t launch_ext_layer(hipStream_t s, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                            const uint8_t* lmask, uint32_t begin, uint32_t count);
// Layers first..last in one work-group (offs: the device offsets of the
// sort; with last_dev, last is read from there on the device).
hipError_t launch_ext_tail(hipStream_t s, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                           const uint8_t* lmask, const uint32_t* offs, int32_t first, int32_t last,
                           const uint32_t* last_dev = nullptr);
// Layers L and L+1 (n0 then n1 list entries from `begin`) in one launch;
// fmask = every face's lower-neighbour mask.
hipError_t launch_ext_layer2(hipStream_t s, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                             const uint8_t* lmask, const uint8_t* fmask, uint32_t begin, uint32_t n0, uint32_t n1);
// buildFluidDensity / buildFluidSDF (EulerSolverTool.cu:324-448) on the
// cell-centred grid; scratch holds ncell int32.
hipError_t launch_fluid_density(hipStream_t s, const GridDev& g, const float* phi, const float* solid, float* out);
hipError_t launch_fluid_sdf(hipStream_t s, const GridDev& g, const float* phi, int32_t passes, int32_t* scratch,
                            float* out);
// ---- particles (mgf_kernels_particles.hip); kind 0 cells, 1..3 face grids ----
size_t pg_field_count(const GridDev& g, int kind);
hipError_t launch_p2g(hipStream_t s, const GridDev& g, int kind, int64_t np, const float* pos, const float* val,
                      int span, int off, int alg, float* field, float* weight, int normalize);
hipError_t launch_g2p(hipStream_t s, const GridDev& g, int kind, int64_t np, const float* pos, const float* field,
                      int alg, float* val, int span, int off);
hipError_t launch_particle_markers(hipStream_t s, const GridDev& g, const float* solid, int64_t np, const float* pos,
                                   float* out);
hipError_t launch_axpy_n(hipStream_t s, size_t n, float a, const float* x, float* y);
hipError_t launch_scale_n(hipStream_t s, size_t n, float a, float* x);
hipError_t launch_clamp_particles(hipStream_t s, const GridDev& g, int64_t np, float* pos);
// Particle generation: tile sums of the fluid cells scanned in place
// (gen_tiles(ncell) + 1 words, the last the total), then the emit pass.
uint32_t gen_tiles(size_t ncell);
hipError_t launch_gen_count(hipStream_t s, size_t ncell, const float* fluid, const float* solid, uint32_t* tiles);
hipError_t launch_gen_emit(hipStream_t s, const GridDev& g, const float* fluid, const float* solid,
                           const uint32_t* tiles, int per_cell, uint64_t seed, float* pos);
hipError_t launch_flip_delta(hipStream_t s, size_t n, const float* in, float* delta);
// Particles counting-sorted by the lower corner of their support on the cell
// grid (bin_count(g) bins; counts, cursor: that many u32, start one more,
// tiles bin_tiles(g) + 1, order np; pos_s 3 np and, when vel is given, vel_s
// 3 np: sorted copies), then the LINEAR splat onto grid `kind` in two
// atomic-free passes through acc (p2g_acc_floats(g) floats); val is in bin
// order when val_sorted, else per particle; every node of the field and
// weight is written.
size_t bin_count(const GridDev& g);
uint32_t bin_tiles(const GridDev& g);
size_t p2g_acc_floats(const GridDev& g);
hipError_t launch_bin_particles(hipStream_t s, const GridDev& g, int64_t np, const float* pos, uint32_t* counts,
                                uint32_t* tiles, uint32_t* start, uint32_t* cursor, uint32_t* order, float* pos_s,
                                const float* vel, float* vel_s);
hipError_t launch_p2g_binned(hipStream_t s, const GridDev& g, int kind, const uint32_t* start, const uint32_t* order,
                             const float* pos_s, const float* val, int span, int off, int val_sorted, float* acc,
                             float* field, float* weight, int normalize);
// Fused particle kernels: the particle-velocity update of PIC / FLIP / MIX
// (delta = NULL for PIC); the first velocity sample of a substep loop with
// its max |v|; one RK substep (resample: the next substep's velocities and
// max |v|).
hipError_t launch_transfer_vel(hipStream_t s, const GridDev& g, int64_t np, const float* pos,
                               const float* const in[3], const float* const delta[3], int kind, int alg, float keep,
                               float mixing, float* vel);
hipError_t launch_sample_absmax(hipStream_t s, const GridDev& g, Vel V, int64_t np, const float* pos, float* out,
                                uint32_t* vmax_bits);
hipError_t launch_rk_substep(hipStream_t s, const GridDev& g, Vel V, int64_t np, float* pos, float* v1, int order,
                             float step, float half, float three_q, float c1, float c2, float c3, int resample,
                             uint32_t* vmax_bits);
hipError_t launch_particle_markers_u8(hipStream_t s, const GridDev& g, const uint8_t* solid_mask, int64_t np,
                                      const float* pos, uint8_t* markers);
// The particle advection loop (mgf_particles.cpp): scratch holds at least 3 n
// floats, vmax one word.
mgf_status particles_advect_run(hipStream_t st, const GridDev& g, Vel V, int64_t n, float* pos, float dt, float cfl,
                                int order, int fix, float* scratch, uint32_t* vmax);
mgf_status psolver_velocity(mgf_psolver s, hipStream_t st, const float* const in[3], const float* const vel[3],
                            float dt, int order, int fix, float* const out[3]);
// reuse_bins: the particles have not moved since the solver's last splat
// (the fields of one library step).
mgf_status psolver_scalar(mgf_psolver s, hipStream_t st, const float* in, float* out, bool reuse_bins);
mgf_status psolver_vector(mgf_psolver s, hipStream_t st, const float* const in[3], float* const out[3],
                          bool reuse_bins);
// Boundaries.
hipError_t launch_resample_solid(hipStream_t s, const GridDev& g, const float* sdf0,
                                 const float t[3], float* out);
hipError_t launch_solids_marker(hipStream_t s, size_t n, const float* sdf, int index1,
                                float* marker);
hipError_t launch_solid_mask(hipStream_t s, size_t n, const float* total, uint8_t* mask);
hipError_t launch_solids_vel(hipStream_t s, const GridDev& g, const float* total,
                             const float* vel3, int nsolids, float* us, float* vs, float* ws);
hipError_t launch_fluid_box(hipStream_t s, const GridDev& g, const float mn[3],
                            const float mx[3], float* phi);
hipError_t launch_u8_to_f32(hipStream_t s, const uint8_t* a, float* b, size_t n);
hipError_t launch_f64_to_f32(hipStream_t s, const double* a, float* b, size_t n);

}  // namespace mgf
