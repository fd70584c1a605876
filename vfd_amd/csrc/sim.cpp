// sim.cpp - host driver of the MI355X DFSPH step and the C ABI declared in include/vfd.h.
//
// Replaces DFSPHImplementation (VFD/Source/Simulation/DFSPH/DFSPHImplementation.h:23-137, .cu:14-839) and the
// DFSPHSimulation wrapper (DFSPHSimulator.cpp). The reference launches ~25 + 3*it_J + 13*it_CG kernels per step, each
// followed by cudaDeviceSynchronize and scalar device-to-host copies. Here a step is a fixed sequence of launches on one
// stream; the only host synchronisations are the grid-bounds read of the neighbour search, one poll of the solver state
// every few Jacobi/CG iterations, and the end-of-step state read.
//
// Compiled with -ffp-contract=off: the host float arithmetic that derives the kernel table, grid and constants is the
// reference's expression order in IEEE single precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vfd.h"
#include "common.h"
#include "grid.h"
#include "kernels.h"
#include "mesh.h"

// run-time knob defaults (vfd_set_tuning); vfd_amd/build.py passes the values chosen by vfd_amd/autotune.py
#ifndef VFD_DEFAULT_USE_GRAPHS
#define VFD_DEFAULT_USE_GRAPHS 1
#endif
#ifndef VFD_DEFAULT_CHECK_INTERVAL
#define VFD_DEFAULT_CHECK_INTERVAL 4
#endif
#ifndef VFD_DEFAULT_PREDICT_ITERATIONS
#define VFD_DEFAULT_PREDICT_ITERATIONS 0
#endif
#ifndef VFD_DEFAULT_HALO_LIMIT
#define VFD_DEFAULT_HALO_LIMIT vfd::kHaloMax
#endif

namespace vfd {

static thread_local std::string g_lastError;

static int fail(int code, const std::string& msg) {
    g_lastError = msg;
    return code;
}
int set_error(int code, const std::string& msg) { return fail(code, msg); }  // for the other host sources

// Set by vfd_step_no_alloc for the duration of its call on this thread: a device buffer that would have to be
// allocated or grown fails the call instead, and the step takes its grid and curvature-table scratch from the
// library's static scratch.
static thread_local bool t_noAlloc = false;
struct NoAllocScope {
    NoAllocScope() { t_noAlloc = true; }
    ~NoAllocScope() { t_noAlloc = false; }
};

#define VFD_HIP(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(VFD_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define VFD_TRY(expr)            \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != VFD_OK) return rc_; \
    } while (0)

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        if (count <= n && p) return VFD_OK;
        if (t_noAlloc && count != 0)
            return fail(VFD_ERROR_STATE, "vfd_step_no_alloc: a device buffer of " +
                                             std::to_string(count * sizeof(T)) +
                                             " bytes would have to be allocated (take one vfd_step, or call "
                                             "vfd_get_particles, after vfd_set_fluid to allocate the state first)");
        release();
        if (count == 0) return VFD_OK;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(VFD_ERROR_OUT_OF_MEMORY, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                                     " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
        return VFD_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// ---- tables (host) -------------------------------------------------------------------------------------------------
// PrecomputedDFSPHCubicKernel::SetRadius (Kernel/DFSPHKernels.h:13-41) in IEEE float.
struct KernelTable {
    float R = 0, R2 = 0, invStep = 0, W0 = 0, K = 0, L = 0;
    std::vector<float> W, G;
    float calcW(float r) const {
        float res = 0.0f;
        const float q = r / R;
        if (q <= 1.0f) {
            if (q <= 0.5f) {
                const float q2 = q * q;
                const float q3 = q2 * q;
                res = K * (6.0f * q3 - 6.0f * q2 + 1.0f);
            } else {
                const float q1 = 1.0f - q;
                const float q3 = q1 * q1 * q1;
                res = K * (2.0f * q3);
            }
        }
        return res;
    }
    // CalculateGradientW({x, 0, 0}).x: |r| = sqrt(fma(0,0,fma(0,0,x*x)))
    float calcGradWx(float x) const {
        const float rl = std::sqrt(std::fma(0.0f, 0.0f, std::fma(0.0f, 0.0f, x * x)));
        const float q = rl / R;
        if (rl > 1.0e-5f && q <= 1.0f) {
            const float gq = x * (1.0f / (rl * R));
            if (q <= 0.5f) {
                const float s = L * q * (3.0f * q - 2.0f);
                return s * gq;
            }
            const float f = 1.0f - q;
            const float s = L * (-f * f);
            return s * gq;
        }
        return 0.0f;
    }
    void set_radius(float radius) {
        const float pi = 3.14159265358979323846f;
        R = radius;
        R2 = R * R;
        const float r3 = R * R * R;
        K = 8.0f / (pi * r3);
        L = 48.0f / (pi * r3);
        W0 = calcW(std::sqrt(std::fma(0.0f, 0.0f, std::fma(0.0f, 0.0f, 0.0f * 0.0f))));
        const float step = R / (float)(kTableResolution - 1);
        invStep = 1.0f / step;
        W.assign(kTableResolution, 0.0f);
        G.assign(kTableResolution + 1, 0.0f);
        for (unsigned i = 0; i < kTableResolution; ++i) {
            const float x = step * (float)i;
            W[i] = calcW(x);
            G[i] = x > 1.0e-9f ? calcGradWx(x) / x : 0.0f;
        }
        G[kTableResolution] = 0.0f;
    }
    // averaged entries: the value GetW/GetGradientW return for table index p
    std::vector<float2> averaged() const {
        std::vector<float2> t(kTableEntries);
        for (unsigned p = 0; p < kTableEntries; ++p) {
            t[p].x = 0.5f * (W[p] + W[p + 1]);
            t[p].y = 0.5f * (G[p] + G[p + 1]);
        }
        return t;
    }
};

// 16,384 unit-sphere points of HaltonVec323.cuh from their closed form (radical inverses in bases 2 and 3),
// rounded to the 15 decimals the reference table holds.
static std::vector<float> make_halton() {
    std::vector<float> h(kHaltonFloats);
    auto radinv = [](unsigned base, unsigned i) {
        const double inv = 1.0 / base;
        double f = inv, r = 0.0;
        while (i) {
            r += f * (double)(i % base);
            i /= base;
            f *= inv;
        }
        return r;
    };
    const double twopi = 6.283185307179586476925286766559;
    for (unsigned i = 0; i < kHaltonFloats / 3; ++i) {
        const double z = 1.0 - 2.0 * radinv(3, i);
        const double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double a = twopi * radinv(2, i);
        h[3 * i + 0] = (float)(std::nearbyint(rxy * std::cos(a) * 1e15) / 1e15);
        h[3 * i + 1] = (float)(std::nearbyint(rxy * std::sin(a) * 1e15) / 1e15);
        h[3 * i + 2] = (float)(std::nearbyint(z * 1e15) / 1e15);
    }
    return h;
}

// 16-point Gauss-Legendre rule on [-1,1] (GaussQuadrature.cpp, p = 30), ascending, by Newton iteration on P_16.
static void make_gauss_legendre(std::vector<float>& xs, std::vector<float>& ws) {
    const int n = 16;
    std::vector<double> x(n), w(n);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    for (int i = 0; i < n / 2; ++i) {
        const double xm = 0.5 * (x[n - 1 - i] - x[i]), wm = 0.5 * (w[i] + w[n - 1 - i]);
        x[i] = -xm;
        x[n - 1 - i] = xm;
        w[i] = w[n - 1 - i] = wm;
    }
    xs.resize(n);
    ws.resize(n);
    for (int i = 0; i < n; ++i) {
        xs[i] = (float)x[i];
        ws[i] = (float)w[i];
    }
}

// ---- scenes -------------------------------------------------------------------------------------------------------
// One captured frame: {x, v, a} per particle in original order (DFSPHParticleSimple), 9 floats each. The data lands
// in pinned host memory by an asynchronous copy that overlaps the following steps; `done` marks its completion.
// When pinned memory cannot be had, the frame is copied synchronously into `pageable`.
struct Frame {
    float* host = nullptr;  // pinned, owned by the simulation's pool
    size_t cap = 0;         // floats in `host`
    size_t n = 0;           // floats of this frame
    hipEvent_t done = nullptr;
    std::vector<float> pageable;
    float maxVel = 0, dt = 0;
    const float* data() const { return host ? host : pageable.data(); }
};

struct Body {
    vfd_boundary_desc desc{};
    MapDev map{};
    DevBuf<float> n0, n1;
};

struct Scene {
    vfd_desc desc{};
    std::vector<float> x0, v0;  // baked particles (SetFluidObjects keeps the host copy, :189-224)
    uint32_t count = 0, base = 0, tileBegin = 0, tileEnd = 0;
    std::vector<Body*> bodies;
    vfd_debug_info dbg{};
    std::vector<Frame> frames;
    SceneState st{};  // host copy
    SceneParams prm{};
    SceneGrid grid{};
    ~Scene() {
        for (Body* b : bodies) delete b;
    }
};

// SetDescription (DFSPHImplementation.cu:318-368)
static void derive_params(const vfd_desc& d, SceneParams& p) {
    const float r = d.ParticleRadius;
    p.r = r;
    p.diam = 2.0f * r;
    p.R = 4.0f * r;
    p.R2 = p.R * p.R;
    p.V = 0.8f * p.diam * p.diam * p.diam;
    p.rho0 = 1000.0f;
    p.m = p.V * p.rho0;
    p.mInv = 1.0f / p.m;
    p.mu = d.Viscosity * p.rho0;
    p.muB = d.BoundaryViscosity * p.rho0;
    p.td = d.TangentialDistanceFactor * p.R;
    p.sigma = d.SurfaceTension;
    p.slope = 74.688796680497925f;
    p.constant = 12.0f;
    p.smoothing = 0.5f;
    p.NR = r * 0.8f;
    p.mcFactor = std::asin(p.NR / r);
    p.gx = d.Gravity[0];
    p.gy = d.Gravity[1];
    p.gz = d.Gravity[2];
    p.divTol = d.MaxDivergenceSolverError;
    p.presTol = d.MaxPressureSolverError;
    p.viscTol = d.MaxViscositySolverError;
    p.dtMin = d.MinTimeStepSize;
    p.dtMax = d.MaxTimeStepSize;
    p.frameLength = d.FrameLength;
    p.minDiv = d.MinDivergenceSolverIterations;
    p.maxDiv = d.MaxDivergenceSolverIterations;
    p.minPres = d.MinPressureSolverIterations;
    p.maxPres = d.MaxPressureSolverIterations;
    p.minVisc = d.MinViscositySolverIterations;
    p.maxVisc = d.MaxViscositySolverIterations;
    p.stPasses = d.SurfaceTensionSmoothPassCount;
    p.temporal = d.TemporalSmoothing ? 1u : 0u;
    p.enableDiv = d.EnableDivergenceSolverError ? 1 : 0;
    p.enableVisc = d.EnableViscositySolver ? 1 : 0;
    p.enableST = d.EnableSurfaceTensionSolver ? 1 : 0;
    p.csdFix = d.CSDFix;
    p.csd = d.CSD;
}

static void set_time_step(SceneState& st, const vfd_desc& d, float dt) {
    st.dt = dt;
    st.dt2 = dt * dt;
    st.dtInv = 1.0f / dt;
    st.dt2Inv = 1.0f / st.dt2;
    // SurfaceTensionSampleCount from dt (the reference sets it only in ComputeTimeStepSize, :413-420)
    st.S = d.CSDFix > 0 ? (uint32_t)d.CSDFix : (uint32_t)(int32_t)((float)d.CSD * dt);
}

static bool valid_desc(const vfd_desc& d, std::string& why) {
    if (!(d.ParticleRadius > 0.0f) || !std::isfinite(d.ParticleRadius)) {
        why = "ParticleRadius must be positive and finite";
        return false;
    }
    if (!(d.TimeStepSize > 0.0f) || !(d.MinTimeStepSize > 0.0f) || !(d.MaxTimeStepSize > 0.0f)) {
        why = "time step sizes must be positive";
        return false;
    }
    return true;
}

// ---- simulation ---------------------------------------------------------------------------------------------------
class Simulation {
public:
    std::vector<Scene*> sc;
    hipStream_t stream = nullptr;
    bool capture = true;
    bool committed = false;
    int device = 0;

    // ---- threads (DFSPHSimulator.cpp:156-166 runs Simulate on a ThreadPool worker while the editor's UI thread
    //      polls the state, the debug info and the frames) ----------------------------------------------------------
    // hostMu guards the host data those polls read: each scene's debug info, frame list, description, time step and
    // particle count. A step takes it only for its host-side bookkeeping after the GPU work has drained.
    mutable std::mutex hostMu;
    std::atomic<int> simState{VFD_STATE_NONE};
    // set for the duration of every driving call (and of the asynchronous worker): a second driver fails instead of
    // racing the first
    std::atomic<bool> busy{false};
    std::thread worker;
    std::mutex workerMu;  // serialises joins of `worker`
    int workerStatus = VFD_OK;
    std::string workerError;
    // DFSPHImplementation::Simulate's frame loop (:47-58); the caller has restored the baked particles
    int simulate_frames() {
        simState = VFD_STATE_SIMULATING;
        for (;;) {
            bool done = true;
            {
                std::lock_guard<std::mutex> lk(hostMu);
                for (Scene* s : sc)
                    if (s->count && s->dbg.FrameIndex < s->desc.FrameCount) done = false;
            }
            if (done) break;
            const int rc = step(VFD_PHASE_POSITION);
            if (rc != VFD_OK) {
                simState = VFD_STATE_NONE;
                return rc;
            }
        }
        simState = VFD_STATE_READY;
        return VFD_OK;
    }
    int join_worker() {
        std::lock_guard<std::mutex> lk(workerMu);
        if (worker.joinable()) worker.join();
        return workerStatus;
    }

    KernelTable kt;
    DevBuf<float2> dTab;
    DevBuf<float> dHalton, dGlx, dGlw;
    DevBuf<float> dMcTab;
    DevBuf<uint32_t> dMcOff;
    std::vector<float> mcTab;     // host copies of this step's tables (kept until the next step's upload)
    std::vector<uint32_t> mcOff;
    // cos(acos(1 - 2 count / S) + MonteCarloFactor) for count = 0..S of every surface-tension scene, in float with the
    // host's libm, the expression of ComputeSurfaceTensionClassificationKernel (DFSPHKernels.cu:950-953); the
    // classification kernel reads it instead of evaluating cos and acos on the device
    // Whether the curvature tables of every surface-tension scene fit the static scratch at any time step the
    // descriptions allow (dt <= max(TimeStepSize, MaxTimeStepSize)), so a no-alloc step cannot fail half-way.
    bool mc_tables_fit_static() const {
        if (sc.size() > kStaticMcScenes) return false;
        double total = 1.0;
        for (const Scene* s : sc) {
            if (!s->desc.EnableSurfaceTensionSolver) continue;
            const double dtMax = std::max((double)s->desc.TimeStepSize, (double)s->desc.MaxTimeStepSize);
            const double S = s->desc.CSDFix > 0 ? (double)s->desc.CSDFix : std::floor((double)s->desc.CSD * dtMax) + 1.0;
            total += S + 1.0;
        }
        return total <= (double)kStaticMcFloats;
    }
    int upload_mc_tables(Fields& F) {
        mcOff.assign(sc.size(), 0u);
        mcTab.clear();
        for (size_t k = 0; k < sc.size(); ++k) {
            const Scene& s = *sc[k];
            mcOff[k] = (uint32_t)mcTab.size();
            if (!s.desc.EnableSurfaceTensionSolver || s.count == 0) continue;
            const uint32_t S = s.st.S;
            for (uint32_t count = 0; count <= S; ++count)
                mcTab.push_back(std::cos(std::acos(1.0f - 2.0f * ((float)count / (float)S)) + s.prm.mcFactor));
        }
        if (mcTab.empty()) mcTab.push_back(0.0f);
        float* tab = nullptr;
        uint32_t* off = nullptr;
        if (t_noAlloc) {  // sizes checked by vfd_step_no_alloc before the step began (mc_table_fits)
            if (mcTab.size() > kStaticMcFloats || mcOff.size() > kStaticMcScenes)
                return fail(VFD_ERROR_NOT_SUPPORTED, "vfd_step_no_alloc: curvature tables exceed the static scratch");
            tab = reinterpret_cast<float*>(staticScratch + kStaticMcTabAt);
            off = staticScratch + kStaticMcOffAt;
        } else {
            VFD_TRY(dMcTab.alloc(mcTab.size()));
            VFD_TRY(dMcOff.alloc(mcOff.size()));
            tab = dMcTab.p;
            off = dMcOff.p;
        }
        VFD_HIP(hipMemcpyAsync(tab, mcTab.data(), mcTab.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        VFD_HIP(hipMemcpyAsync(off, mcOff.data(), mcOff.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        F.mcTab = tab;
        F.mcOff = off;
        return VFD_OK;
    }

    uint32_t tiles = 0, cap = 0;
    int cur = 0;
    DevBuf<SceneParams> dP;
    DevBuf<uint32_t> dTileScene;
    DevBuf<SceneState> dState;
    DevBuf<SceneGrid> dGrid;
    DevBuf<MapDev> dMaps;
    DevBuf<int> dMM, dMMPart;
    DevBuf<float4> pos, posNext, vel[2], dv[2], acc, ap, nMC, nS, xb, cgX, cgR, cgP, cgZ, cgAp;
    DevBuf<float> dk[2], density, factor, densAdv, kappa, kappaV, residuum, kMC, kS, vb, minv;
    DevBuf<uint32_t> id[2], nbrCount, nbrList, cellOf, rank, tmpId, cellCount, cellStart, scanScratch, live;
    DevBuf<double> partials;
    DevBuf<float> partialsF;
    DevBuf<unsigned long long> nbStats;
    DevBuf<uint32_t> haloCount, haloIds, haloStats;
    DevBuf<uint32_t> nbrLocal;

    // ---- adjoint recording (one scene): per step, everything the reverse needs, in that step's slot order -----
    struct DevTape {
        DevBuf<float4> x, v0, v1, y, v2;
        DevBuf<float> alpha, D, rho, rhoadv, kvEnd, kpEnd;
        DevBuf<uint32_t> id, count, list;
        // boundary pass: contact samples and push-outs with their derivatives (adj_launch_boundary_tape) and
        // flags [contact, push, out of scope]
        DevBuf<float4> con, push;
        DevBuf<uint32_t> flags;
        // surface tension: Monte-Carlo normals and curvatures of the step (scenes that run it)
        DevBuf<float4> nMC;
        DevBuf<float> kMC, dK0;  // dK0: the blended curvature carried into the step (temporal smoothing)
        std::deque<DevBuf<float>> kv, rv, kp, rp;  // per Jacobi iteration: kappa^k, residual^k
        std::deque<DevBuf<float4>> av, ap;         // per Jacobi iteration: pressure acceleration a^k
        uint32_t n = 0;                            // slots recorded (every scene's)
        bool st = false, stTemporal = false;       // some scene ran surface tension, some with temporal smoothing
        // per scene: time steps before and after the step's CFL update, Jacobi iteration counts, bodies, switches
        struct SceneRec {
            float h = 0, H = 0;
            uint32_t Kd = 0, Kp = 0, bodies = 0, stPasses = 0;
            bool visc = false, boundaryViscosity = false, st = false, stTemporal = false, atCap = false;
        };
        std::vector<SceneRec> scn;
    };
    DevBuf<float4> tapeXpre;  // positions the boundary pass of a recorded step starts from
    // recorded steps: the first tapeCount entries of tapes; the buffers are kept and reused by later recordings
    std::deque<std::unique_ptr<DevTape>> tapes;
    size_t tapeCount = 0;
    bool taping = false;
    std::string tapeError;
    template <class U> int tape_copy(DevBuf<U>& dst, const U* src, size_t count) {
        VFD_TRY(dst.alloc(count));
        VFD_HIP(hipMemcpyAsync(dst.p, src, count * sizeof(U), hipMemcpyDeviceToDevice, stream));
        return VFD_OK;
    }
    template <class U> int tape_slot(std::deque<DevBuf<U>>& q, uint32_t k, const U* src, size_t count) {
        while (q.size() <= k) q.emplace_back();
        return tape_copy(q[k], src, count);
    }
    // the end of Jacobi iteration k in one launch: a^k (from the acceleration pass), the residual of iteration k and
    // kappa^(k+1) (the updated kappa, the next iteration's start)
    int tape_iteration(std::deque<DevBuf<float>>& kq, std::deque<DevBuf<float4>>& aq, std::deque<DevBuf<float>>& rq,
                       uint32_t k, uint32_t n, const float* kap, const float4* a, const float* r) {
        while (kq.size() <= k + 1) kq.emplace_back();
        while (aq.size() <= k) aq.emplace_back();
        while (rq.size() <= k) rq.emplace_back();
        VFD_TRY(kq[k + 1].alloc(n));
        VFD_TRY(aq[k].alloc(n));
        VFD_TRY(rq[k].alloc(n));
        adj_launch_copy3(stream, n, kap, kq[k + 1].p, a, aq[k].p, r, rq[k].p);
        return VFD_OK;
    }
    uint32_t haloLimit = VFD_DEFAULT_HALO_LIMIT;  // largest halo staged in LDS (tuning)
    uint32_t cellCap = 0;
    hipEvent_t ev[8] = {};
    bool eventsMade = false;
    // solver polling: live[which] is copied into one of two pinned slots after each chunk of iterations
    uint32_t* hostLive = nullptr;  // [2]
    hipEvent_t pollEv[2] = {};
    uint32_t pollChunk = VFD_DEFAULT_CHECK_INTERVAL;  // iterations enqueued between convergence checks
    // One chunk of pollChunk iterations of each solver loop, captured as a HIP graph on first use. The launches'
    // arguments differ only with the double-buffered particle arrays (index cur), so there is one graph per
    // (solver, cur); they are dropped whenever buffers are reallocated or the chunk length changes.
    bool useGraphs = VFD_DEFAULT_USE_GRAPHS != 0;
    hipGraphExec_t loopGraph[3][2] = {};
    bool graphsStale = false;            // dropped during a no-alloc step; destroyed by the next call that may free
    uint32_t* staticScratch = nullptr;   // this device's static scratch (vfd_step_no_alloc), looked up at init
    // Iterations each solver loop ran in the previous step (most over the scenes; Jacobi divergence, Jacobi
    // pressure, CG). With predictIterations the loop's first chunk is one iteration shorter than that, so a solve
    // that needs as many iterations as last time ends with one empty iteration in flight instead of a whole chunk.
    // Off by default: on the 2M dam break it saved 0.1 ms of empty launches per step but only 0.2% of the step
    // (with SolverCheckInterval 1 or 2), within run-to-run noise.
    bool predictIterations = VFD_DEFAULT_PREDICT_ITERATIONS != 0;
    uint32_t lastIterations[3] = {0, 0, 0};
    hipStream_t ownStream = nullptr;  // default stream (capture needs a stream other than the legacy null stream)
    // frame output: frames are packed in original order on the simulation's stream into one of two device staging
    // buffers and copied to pinned host memory on copyStream, so that the copy overlaps the next steps
    hipStream_t copyStream = nullptr;
    DevBuf<float> frameStage[2];
    hipEvent_t stageFree[2] = {}, stagePacked[2] = {};
    bool stageUsed[2] = {false, false};
    uint32_t stageNext = 0;
    std::vector<std::pair<float*, size_t>> pinnedPool;  // pinned buffers of released frames, for reuse
    std::vector<hipEvent_t> frameEvents;                // events of released frames, for reuse
    int capture_frame(uint32_t s, Frame& fr) {
        const Scene& S = *sc[s];
        fr.n = (size_t)9 * S.count;
        // pinned destination: the smallest pooled buffer that fits, else a new one
        size_t best = pinnedPool.size();
        for (size_t k = 0; k < pinnedPool.size(); ++k)
            if (pinnedPool[k].second >= fr.n && (best == pinnedPool.size() || pinnedPool[k].second < pinnedPool[best].second))
                best = k;
        if (best < pinnedPool.size()) {
            fr.host = pinnedPool[best].first;
            fr.cap = pinnedPool[best].second;
            pinnedPool.erase(pinnedPool.begin() + (long)best);
        } else if (hipHostMalloc((void**)&fr.host, fr.n * sizeof(float), hipHostMallocDefault) == hipSuccess) {
            fr.cap = fr.n;
        } else {
            (void)hipGetLastError();
            fr.host = nullptr;
        }
        const uint32_t b = stageNext;
        stageNext ^= 1u;
        VFD_TRY(frameStage[b].alloc(fr.n));
        if (stageUsed[b]) VFD_HIP(hipStreamWaitEvent(stream, stageFree[b], 0));  // its previous copy has finished
        launch_frame_pack(stream, S.base, S.count, id[cur].p, posNext.p, vel[cur].p, acc.p, frameStage[b].p);
        VFD_HIP(hipGetLastError());
        if (!fr.host) {
            fr.pageable.resize(fr.n);
            VFD_HIP(hipMemcpyAsync(fr.pageable.data(), frameStage[b].p, fr.n * sizeof(float), hipMemcpyDeviceToHost,
                                   stream));
            VFD_HIP(hipStreamSynchronize(stream));
            return VFD_OK;
        }
        if (frameEvents.empty()) {
            hipEvent_t e;
            VFD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            frameEvents.push_back(e);
        }
        fr.done = frameEvents.back();
        frameEvents.pop_back();
        VFD_HIP(hipEventRecord(stagePacked[b], stream));
        VFD_HIP(hipStreamWaitEvent(copyStream, stagePacked[b], 0));
        VFD_HIP(hipMemcpyAsync(fr.host, frameStage[b].p, fr.n * sizeof(float), hipMemcpyDeviceToHost, copyStream));
        VFD_HIP(hipEventRecord(fr.done, copyStream));
        VFD_HIP(hipEventRecord(stageFree[b], copyStream));
        stageUsed[b] = true;
        return VFD_OK;
    }
    // frames of a scene go back to the pools once their copies have completed
    void release_frames(Scene& S) {
        for (Frame& f : S.frames) {
            if (f.done) {
                (void)hipEventSynchronize(f.done);
                frameEvents.push_back(f.done);
            }
            if (f.host) pinnedPool.emplace_back(f.host, f.cap);
        }
        S.frames.clear();
    }
    void drop_graphs() {
        if (t_noAlloc) {  // hipGraphExecDestroy frees: leave it to the next call that may, and use no graph till then
            graphsStale = true;
            return;
        }
        graphsStale = false;
        for (auto& row : loopGraph)
            for (auto& g : row)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
    }

    // per-kernel timing: an event pair around each launch, resolved at the end of a step
    struct KStat {
        std::string name;
        uint32_t launches = 0;
        double ms = 0.0;
    };
    bool profiling = false;
    std::vector<KStat> kstats;
    std::vector<hipEvent_t> evPool;
    struct Pending {
        size_t stat, event;  // stat index, first event of the pair
        int solver;          // -1, or the solver loop (0 divergence, 1 pressure, 2 viscosity) the launch belongs to
        uint32_t iter;       // its iteration index within that loop
    };
    std::vector<Pending> pending;
    size_t evUsed = 0;

    ~Simulation() {
        join_worker();
        if (eventsMade) {
            for (auto& e : ev) (void)hipEventDestroy(e);
            for (auto& e : pollEv) (void)hipEventDestroy(e);
        }
        if (hostLive) (void)hipHostFree(hostLive);
        drop_graphs();
        if (copyStream) (void)hipStreamSynchronize(copyStream);
        for (Scene* s : sc) release_frames(*s);
        for (auto& pb : pinnedPool) (void)hipHostFree(pb.first);
        for (auto& e : frameEvents) (void)hipEventDestroy(e);
        for (int b = 0; b < 2; ++b) {
            if (stageFree[b]) (void)hipEventDestroy(stageFree[b]);
            if (stagePacked[b]) (void)hipEventDestroy(stagePacked[b]);
        }
        if (copyStream) (void)hipStreamDestroy(copyStream);
        if (ownStream) (void)hipStreamDestroy(ownStream);
        for (auto& e : evPool) (void)hipEventDestroy(e);
        for (Scene* s : sc) delete s;
    }

    size_t stat_index(const char* name) {
        for (size_t i = 0; i < kstats.size(); ++i)
            if (kstats[i].name == name) return i;
        kstats.push_back(KStat{name, 0, 0.0});
        return kstats.size() - 1;
    }
    // Runs `fn` (which enqueues launches on the stream) between two recorded events when profiling. Launches of a
    // solver loop are attributed at the end of the step: those past the last iteration any scene ran (issued ahead of
    // the convergence check and skipped on the device) are counted under "<name>_idle".
    template <class Fn> int timed(const char* name, Fn fn, int solver = -1, uint32_t iter = 0) {
        if (!profiling) {
            fn();
            return VFD_OK;
        }
        while (evPool.size() < evUsed + 2) {
            hipEvent_t e;
            VFD_HIP(hipEventCreate(&e));
            evPool.push_back(e);
        }
        const size_t first = evUsed;
        evUsed += 2;
        VFD_HIP(hipEventRecord(evPool[first], stream));
        fn();
        VFD_HIP(hipEventRecord(evPool[first + 1], stream));
        pending.push_back(Pending{stat_index(name), first, solver, iter});
        return VFD_OK;
    }
    int resolve_profile() {
        if (pending.empty()) return VFD_OK;
        VFD_HIP(hipStreamSynchronize(stream));
        uint32_t ran[3] = {0, 0, 0};
        for (Scene* s : sc) {
            if (!s->count) continue;
            ran[0] = std::max(ran[0], s->st.div.it);
            ran[1] = std::max(ran[1], s->st.pres.it);
            ran[2] = std::max(ran[2], s->st.cg.loops);
        }
        for (const auto& pr : pending) {
            float ms = 0.0f;
            VFD_HIP(hipEventElapsedTime(&ms, evPool[pr.event], evPool[pr.event + 1]));
            size_t k = pr.stat;
            if (pr.solver >= 0 && pr.iter >= ran[pr.solver]) k = stat_index((kstats[k].name + "_idle").c_str());
            kstats[k].launches++;
            kstats[k].ms += ms;
        }
        pending.clear();
        evUsed = 0;
        return VFD_OK;
    }

    int init(const vfd_desc* descs, uint32_t n) {
        if (n == 0) return fail(VFD_ERROR_INVALID_ARGUMENT, "scene_count must be positive");
        for (uint32_t i = 0; i < n; ++i) {
            std::string why;
            if (!valid_desc(descs[i], why)) return fail(VFD_ERROR_INVALID_ARGUMENT, why);
            if (descs[i].ParticleRadius != descs[0].ParticleRadius)
                return fail(VFD_ERROR_NOT_SUPPORTED, "all scenes of a batch must use the same ParticleRadius");
        }
        int dev = 0, cus = 0;
        VFD_HIP(hipGetDevice(&dev));
        device = dev;
        VFD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        numCU = cus > 0 ? (uint32_t)cus : 256u;
        if (configure_kernels() != 0) return fail(VFD_ERROR_HIP, "hipFuncSetAttribute (dynamic LDS) failed");
        if (static_scratch_address(dev, &staticScratch) != 0)
            return fail(VFD_ERROR_HIP, "hipGetSymbolAddress of the static scratch failed");
        for (uint32_t i = 0; i < n; ++i) {
            Scene* s = new Scene();
            s->desc = descs[i];
            derive_params(descs[i], s->prm);
            set_time_step(s->st, descs[i], descs[i].TimeStepSize);
            sc.push_back(s);
        }
        VFD_TRY(build_tables(descs[0].ParticleRadius));
        for (auto& e : ev) VFD_HIP(hipEventCreate(&e));
        for (auto& e : pollEv) VFD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        eventsMade = true;
        VFD_HIP(hipHostMalloc((void**)&hostLive, 2 * sizeof(uint32_t), hipHostMallocDefault));
        VFD_HIP(hipStreamCreateWithFlags(&ownStream, hipStreamNonBlocking));
        stream = ownStream;
        VFD_HIP(hipStreamCreateWithFlags(&copyStream, hipStreamNonBlocking));
        for (int b = 0; b < 2; ++b) {
            VFD_HIP(hipEventCreateWithFlags(&stageFree[b], hipEventDisableTiming));
            VFD_HIP(hipEventCreateWithFlags(&stagePacked[b], hipEventDisableTiming));
        }
        return VFD_OK;
    }

    int build_tables(float r) {
        drop_graphs();  // the table address is a captured launch argument
        kt.set_radius(4.0f * r);
        const std::vector<float2> avg = kt.averaged();
        VFD_TRY(dTab.alloc(avg.size()));
        VFD_HIP(hipMemcpy(dTab.p, avg.data(), avg.size() * sizeof(float2), hipMemcpyHostToDevice));
        if (!dHalton.p) {
            const std::vector<float> h = make_halton();
            VFD_TRY(dHalton.alloc(h.size()));
            VFD_HIP(hipMemcpy(dHalton.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
            std::vector<float> gx, gw;
            make_gauss_legendre(gx, gw);
            VFD_TRY(dGlx.alloc(16));
            VFD_TRY(dGlw.alloc(16));
            VFD_HIP(hipMemcpy(dGlx.p, gx.data(), 16 * sizeof(float), hipMemcpyHostToDevice));
            VFD_HIP(hipMemcpy(dGlw.p, gw.data(), 16 * sizeof(float), hipMemcpyHostToDevice));
        }
        return VFD_OK;
    }

    KTabHost ktab() const { return KTabHost{dTab.p, kt.R, kt.R2, kt.invStep, kt.W0}; }

    int set_description(uint32_t s, const vfd_desc& d) {
        std::string why;
        if (!valid_desc(d, why)) return fail(VFD_ERROR_INVALID_ARGUMENT, why);
        if (d.ParticleRadius != sc[s]->desc.ParticleRadius) {
            if (sc.size() > 1) return fail(VFD_ERROR_NOT_SUPPORTED, "a batch cannot change ParticleRadius");
            if (!sc[s]->bodies.empty())
                return fail(VFD_ERROR_STATE, "ParticleRadius cannot change after boundaries are built");
            VFD_TRY(build_tables(d.ParticleRadius));
        }
        {
            std::lock_guard<std::mutex> lk(hostMu);
            sc[s]->desc = d;
        }
        const SceneParams old = sc[s]->prm;
        derive_params(d, sc[s]->prm);
        sc[s]->prm.base = old.base;
        sc[s]->prm.count = old.count;
        sc[s]->prm.tileBegin = old.tileBegin;
        sc[s]->prm.tileEnd = old.tileEnd;
        sc[s]->prm.bodyCount = old.bodyCount;
        {
            std::lock_guard<std::mutex> lk(hostMu);
            set_time_step(sc[s]->st, d, d.TimeStepSize);  // SetDescription resets the time step (:352-356)
        }
        if (committed) {
            VFD_HIP(hipMemcpyAsync(dP.p + s, &sc[s]->prm, sizeof(SceneParams), hipMemcpyHostToDevice, stream));
            VFD_TRY(push_state());
        }
        return VFD_OK;
    }

    int set_fluid(uint32_t s, const float* x, const float* v, uint32_t n) {
        if (n && !x) return fail(VFD_ERROR_INVALID_ARGUMENT, "positions must not be NULL");
        for (size_t i = 0; i < (size_t)3 * n; ++i)
            if (!std::isfinite(x[i]) || (v && !std::isfinite(v[i])))
                return fail(VFD_ERROR_INVALID_ARGUMENT, "non-finite particle position or velocity");
        Scene& S = *sc[s];
        S.x0.assign(x, x + (size_t)3 * n);
        if (v)
            S.v0.assign(v, v + (size_t)3 * n);
        else
            S.v0.assign((size_t)3 * n, 0.0f);
        {
            std::lock_guard<std::mutex> lk(hostMu);
            S.count = n;
            release_frames(S);
            S.dbg = vfd_debug_info{};
        }
        simState = VFD_STATE_NONE;  // SetFluidObjects (:177)
        committed = false;
        tapeCount = 0;  // recorded steps belong to the particles they were recorded with
        return VFD_OK;
    }

    // RigidBody ctor (RigidBody.cu:10-73): volume map domain, resolution and both fields, built on the GPU.
    int add_boundary(uint32_t s, const vfd_boundary_desc& b) {
        if (b.Shape < VFD_SHAPE_BOX || b.Shape > VFD_SHAPE_CYLINDER)
            return fail(VFD_ERROR_INVALID_ARGUMENT, "unknown boundary shape");
        if (b.Shape == VFD_SHAPE_PLANE && !b.HasDomain)
            return fail(VFD_ERROR_INVALID_ARGUMENT, "a plane boundary needs an explicit map domain");
        const float r = sc[s]->desc.ParticleRadius, R = kt.R;
        const float tol = b.Padding - r;
        ShapeDev sh{};
        sh.shape = b.Shape;
        for (int a = 0; a < 3; ++a) {
            sh.c[a] = b.Center[a];
            sh.h[a] = b.HalfExtents[a];
            sh.n[a] = b.Normal[a];
        }
        sh.radius = b.Radius;
        sh.sign = b.Inverted ? -1.0f : 1.0f;
        sh.tol = tol;
        float lo[3], hi[3];
        if (b.HasDomain) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = b.DomainMin[a];
                hi[a] = b.DomainMax[a];
            }
        } else {
            float e[3];
            if (b.Shape == VFD_SHAPE_SPHERE)
                e[0] = e[1] = e[2] = b.Radius;
            else if (b.Shape == VFD_SHAPE_CYLINDER) {
                e[0] = e[2] = b.Radius;
                e[1] = b.HalfExtents[1];
            } else
                for (int a = 0; a < 3; ++a) e[a] = b.HalfExtents[a];
            const float grow = 8.0f * R + tol;  // RigidBody.cu:35-37
            for (int a = 0; a < 3; ++a) {
                lo[a] = (b.Center[a] - e[a]) - grow;
                hi[a] = (b.Center[a] + e[a]) + grow;
            }
        }
        return add_body(s, b.CollisionMapResolution, lo, hi, b,
                        [&](const MapDev& m, float* n0) { launch_map_field0(stream, sh, m, n0); });
    }

    // RigidBody ctor for a mesh (RigidBody.cu:10-73): transformed vertices, domain = their bounding box grown by
    // 8R + tol, field 0 = sign * (MeshDistance::SignedDistance - tol) from the GPU distance kernel.
    int add_boundary_mesh(uint32_t s, const vfd_mesh_boundary_desc& b, const vfd_mesh& mesh) {
        MeshPrep prep;
        VFD_TRY(prepare_mesh(mesh, b.Transform, prep));
        const float r = sc[s]->desc.ParticleRadius, R = kt.R;
        const float tol = b.Padding - r;
        const float grow = 8.0f * R + tol;
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = prep.lo[a] - grow;
            hi[a] = prep.hi[a] + grow;
        }
        DevBuf<float> tri, nrm, sph;
        VFD_TRY(tri.alloc(prep.tri.size()));
        VFD_TRY(nrm.alloc(prep.normals.size()));
        VFD_TRY(sph.alloc(prep.spheres.size()));
        // destroyed before the buffers and `prep`: every return, early error ones included, drains the uploads and
        // the distance kernel that read them
        struct DrainGuard {
            hipStream_t s;
            ~DrainGuard() { (void)hipStreamSynchronize(s); }
        } drain{stream};
        VFD_HIP(hipMemcpyAsync(sph.p, prep.spheres.data(), prep.spheres.size() * sizeof(float), hipMemcpyHostToDevice,
                               stream));
        VFD_HIP(hipMemcpyAsync(tri.p, prep.tri.data(), prep.tri.size() * sizeof(float), hipMemcpyHostToDevice,
                               stream));
        VFD_HIP(hipMemcpyAsync(nrm.p, prep.normals.data(), prep.normals.size() * sizeof(float),
                               hipMemcpyHostToDevice, stream));
        const MeshDev md{(const float4*)tri.p, nrm.p, (const float4*)sph.p, mesh.TriangleCount, prep.firstPos};
        vfd_boundary_desc d{};
        d.Shape = -1;  // mesh
        d.Inverted = b.Inverted;
        d.Padding = b.Padding;
        for (int a = 0; a < 3; ++a) d.CollisionMapResolution[a] = b.CollisionMapResolution[a];
        const float sign = b.Inverted ? -1.0f : 1.0f;
        return add_body(s, b.CollisionMapResolution, lo, hi, d,
                        [&](const MapDev& m, float* n0) { launch_mesh_field0(stream, md, m, sign, tol, n0); });
    }

    // Common tail of both: map geometry (SDF.cu:8-14), node storage, field 0 by `field0`, field 1 (the boundary
    // volume) and the body's registration.
    template <class Field0>
    int add_body(uint32_t s, const uint32_t res[3], const float lo[3], const float hi[3], const vfd_boundary_desc& b,
                 Field0 field0) {
        Scene& S = *sc[s];
        if (S.bodies.size() >= kMaxBodies)
            return fail(VFD_ERROR_NOT_SUPPORTED, "at most " + std::to_string(kMaxBodies) + " boundary bodies per scene");
        for (int a = 0; a < 3; ++a)
            if (res[a] == 0) return fail(VFD_ERROR_INVALID_ARGUMENT, "CollisionMapResolution must be positive");
        const uint64_t cells = (uint64_t)res[0] * res[1] * res[2];
        if (cells > (1u << 26)) return fail(VFD_ERROR_INVALID_ARGUMENT, "CollisionMapResolution too large");
        Body* body = new Body();
        body->desc = b;
        MapDev& m = body->map;
        for (int a = 0; a < 3; ++a) {
            m.domMin[a] = lo[a];
            m.domMax[a] = hi[a];
            m.res[a] = res[a];
            m.cellSize[a] = (m.domMax[a] - m.domMin[a]) / (float)m.res[a];
            m.cellSizeInv[a] = 1.0f / m.cellSize[a];
        }
        const uint32_t nx = m.res[0], ny = m.res[1], nz = m.res[2];
        m.nodeCount = (nx + 1) * (ny + 1) * (nz + 1) +
                      2 * (nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz);
        int rc = body->n0.alloc(m.nodeCount);
        if (rc == VFD_OK) rc = body->n1.alloc(m.nodeCount);
        if (rc != VFD_OK) {
            delete body;
            return rc;
        }
        m.nodes0 = body->n0.p;
        m.nodes1 = body->n1.p;
        field0(m, body->n0.p);
        launch_map_field1(stream, m, body->n0.p, body->n1.p, ktab(), dGlx.p, dGlw.p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            delete body;
            return fail(VFD_ERROR_HIP, std::string("volume map build: ") + hipGetErrorString(e));
        }
        S.bodies.push_back(body);
        S.prm.bodyCount = (uint32_t)S.bodies.size();
        simState = VFD_STATE_NONE;  // SetRigidBodies (:259)
        committed = false;
        return VFD_OK;
    }

    int clear_boundaries(uint32_t s) {
        for (Body* b : sc[s]->bodies) delete b;
        sc[s]->bodies.clear();
        sc[s]->prm.bodyCount = 0;
        simState = VFD_STATE_NONE;
        committed = false;
        return VFD_OK;
    }

    // Allocates the packed particle arrays for the current scenes and uploads the baked particles.
    int commit() {
        drop_graphs();
        tiles = 0;
        for (Scene* s : sc) {
            s->base = tiles * kTile;
            s->tileBegin = tiles;
            tiles += (s->count + kTile - 1) / kTile;
            s->tileEnd = tiles;
            s->prm.base = s->base;
            s->prm.count = s->count;
            s->prm.tileBegin = s->tileBegin;
            s->prm.tileEnd = s->tileEnd;
            s->prm.bodyCount = (uint32_t)s->bodies.size();
        }
        cap = tiles * kTile;
        // a pair of staged-list rows (cap 8-byte entries) is read as one buffer whose byte offsets stay below 2^31
        if ((uint64_t)cap * 2 * sizeof(uint32_t) >= (1ull << 31))
            return fail(VFD_ERROR_NOT_SUPPORTED, "at most 2^28 particle slots in one simulation");
        const uint32_t S = (uint32_t)sc.size();
        std::vector<uint32_t> ts(tiles);
        for (uint32_t k = 0; k < S; ++k)
            for (uint32_t t = sc[k]->tileBegin; t < sc[k]->tileEnd; ++t) ts[t] = k;
        const size_t c = std::max<size_t>(cap, 1);
        VFD_TRY(dP.alloc(S));
        VFD_TRY(dState.alloc(S));
        VFD_TRY(dGrid.alloc(S));
        VFD_TRY(dMaps.alloc((size_t)S * kMaxBodies));
        VFD_TRY(dMM.alloc((size_t)S * 6));
        VFD_TRY(dMMPart.alloc((size_t)std::max<uint32_t>(tiles, 1) * 6));
        VFD_TRY(live.alloc(4));
        VFD_TRY(dTileScene.alloc(std::max<uint32_t>(tiles, 1)));
        DevBuf<float4>* f4[] = {&pos, &posNext, &vel[0], &vel[1], &dv[0], &dv[1], &acc, &ap, &nMC, &nS,
                                &cgX, &cgR, &cgP, &cgZ, &cgAp};
        for (auto* b : f4) VFD_TRY(b->alloc(c));
        DevBuf<float>* f1[] = {&dk[0], &dk[1], &density, &factor, &densAdv, &kappa, &kappaV, &residuum, &kMC, &kS};
        for (auto* b : f1) VFD_TRY(b->alloc(c));
        DevBuf<uint32_t>* u1[] = {&id[0], &id[1], &nbrCount, &cellOf, &rank, &tmpId};
        for (auto* b : u1) VFD_TRY(b->alloc(c));
        VFD_TRY(xb.alloc(c * kMaxBodies));
        VFD_TRY(vb.alloc(c * kMaxBodies));
        VFD_TRY(minv.alloc(c * 9));
        VFD_TRY(nbrList.alloc(c * kMaxNeighbors));
        VFD_TRY(partials.alloc((size_t)std::max<uint32_t>(tiles, 1) * kPartialSlots));
        VFD_TRY(partialsF.alloc(std::max<uint32_t>(tiles, 1)));
        VFD_TRY(nbStats.alloc((size_t)2 * S));
        {
            const Launch L = launch();
            VFD_TRY(haloCount.alloc(std::max<uint32_t>(L.groups, 1)));
            VFD_TRY(haloIds.alloc((size_t)std::max<uint32_t>(L.groups, 1) * kHaloMax));
            VFD_TRY(haloStats.alloc(3));
            VFD_TRY(nbrLocal.alloc(c * kMaxNeighbors));
        }
        if (tiles) VFD_HIP(hipMemcpy(dTileScene.p, ts.data(), tiles * sizeof(uint32_t), hipMemcpyHostToDevice));
        std::vector<MapDev> maps((size_t)S * kMaxBodies);
        for (uint32_t k = 0; k < S; ++k)
            for (size_t b = 0; b < sc[k]->bodies.size(); ++b) maps[(size_t)k * kMaxBodies + b] = sc[k]->bodies[b]->map;
        VFD_HIP(hipMemcpy(dMaps.p, maps.data(), maps.size() * sizeof(MapDev), hipMemcpyHostToDevice));
        // every particle field starts at zero, as SetFluidObjects initialises DFSPHParticle (:198-222)
        VFD_HIP(hipMemset(live.p, 0, 4 * sizeof(uint32_t)));
        VFD_HIP(hipMemset(xb.p, 0, c * kMaxBodies * sizeof(float4)));
        VFD_HIP(hipMemset(vb.p, 0, c * kMaxBodies * sizeof(float)));
        VFD_HIP(hipMemset(minv.p, 0, c * 9 * sizeof(float)));
        VFD_HIP(hipMemset(nbrCount.p, 0, c * sizeof(uint32_t)));
        for (auto* b : f4) VFD_HIP(hipMemset(b->p, 0, c * sizeof(float4)));
        for (auto* b : f1) VFD_HIP(hipMemset(b->p, 0, c * sizeof(float)));
        std::vector<SceneParams> prm(S);
        for (uint32_t k = 0; k < S; ++k) prm[k] = sc[k]->prm;
        VFD_HIP(hipMemcpy(dP.p, prm.data(), S * sizeof(SceneParams), hipMemcpyHostToDevice));
        committed = true;
        VFD_TRY(upload_baked());
        return VFD_OK;
    }

    // Simulate() start (DFSPHImplementation.cu:36-45): baked particles, dv = dk = 0, dt = TimeStepSize.
    int upload_baked() {
        const size_t c = std::max<size_t>(cap, 1);
        std::vector<float4> x(c, make_float4(0, 0, 0, 0)), v(c, make_float4(0, 0, 0, 0));
        std::vector<uint32_t> ids(c, 0);
        for (Scene* s : sc) {
            for (uint32_t i = 0; i < s->count; ++i) {
                x[s->base + i] = make_float4(s->x0[3 * i], s->x0[3 * i + 1], s->x0[3 * i + 2], 0.0f);
                v[s->base + i] = make_float4(s->v0[3 * i], s->v0[3 * i + 1], s->v0[3 * i + 2], 0.0f);
                ids[s->base + i] = i;
            }
            std::lock_guard<std::mutex> lk(hostMu);
            s->dbg = vfd_debug_info{};
            release_frames(*s);
        }
        cur = 0;
        VFD_HIP(hipMemcpy(posNext.p, x.data(), c * sizeof(float4), hipMemcpyHostToDevice));
        VFD_HIP(hipMemcpy(pos.p, x.data(), c * sizeof(float4), hipMemcpyHostToDevice));
        VFD_HIP(hipMemcpy(vel[0].p, v.data(), c * sizeof(float4), hipMemcpyHostToDevice));
        VFD_HIP(hipMemcpy(id[0].p, ids.data(), c * sizeof(uint32_t), hipMemcpyHostToDevice));
        VFD_HIP(hipMemset(dv[0].p, 0, c * sizeof(float4)));
        VFD_HIP(hipMemset(dk[0].p, 0, c * sizeof(float)));
        for (Scene* s : sc) set_time_step(s->st, s->desc, s->desc.TimeStepSize);
        return push_state();
    }

    int push_state() {
        std::vector<SceneState> st(sc.size());
        for (size_t k = 0; k < sc.size(); ++k) st[k] = sc[k]->st;
        VFD_HIP(hipMemcpyAsync(dState.p, st.data(), st.size() * sizeof(SceneState), hipMemcpyHostToDevice, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        return VFD_OK;
    }
    int pull_state() {
        std::vector<SceneState> st(sc.size());
        VFD_HIP(hipMemcpyAsync(st.data(), dState.p, st.size() * sizeof(SceneState), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        for (size_t k = 0; k < sc.size(); ++k) sc[k]->st = st[k];
        return VFD_OK;
    }

    uint32_t numCU = 256;

    Launch launch() const {
        Launch L{};
        L.stream = stream;
        L.P = dP.p;
        L.tileScene = dTileScene.p;
        L.tiles = tiles;
        L.sceneCount = (uint32_t)sc.size();
        L.capacity = cap;
        L.tilesPerXcd = (tiles + 7) / 8;
        const uint32_t per = (L.tilesPerXcd + 3) / 4 * 4;
        L.gridAll = 8 * per;
        const uint32_t p4 = std::max<uint32_t>(8, (numCU * 4) / 8 * 8), p2 = std::max<uint32_t>(8, (numCU * 2) / 8 * 8);
        L.grid4 = std::min(L.gridAll, p4);
        L.grid2 = std::min(L.gridAll, p2);
        L.groups = 8 * (per / kGroupTiles);
        L.gridGroups = std::min(L.groups, std::max<uint32_t>(8, numCU * kGroupsPerCU / 8 * 8));
        L.tg.tiles = tiles;
        L.tg.per = per;
        return L;
    }
    Fields fields() {
        Fields F{};
        F.pos = pos.p;
        F.posNext = posNext.p;
        F.vel = vel[cur].p;
        F.acc = acc.p;
        F.ap = ap.p;
        F.dv = dv[cur].p;
        F.dk = dk[cur].p;
        F.id = id[cur].p;
        F.density = density.p;
        F.factor = factor.p;
        F.densAdv = densAdv.p;
        F.kappa = kappa.p;
        F.kappaV = kappaV.p;
        F.residuum = residuum.p;
        F.nMC = nMC.p;
        F.nS = nS.p;
        F.kMC = kMC.p;
        F.kS = kS.p;
        F.xb = xb.p;
        F.vb = vb.p;
        F.nbrCount = nbrCount.p;
        F.nbrList = nbrList.p;
        F.minv = minv.p;
        F.cgX = cgX.p;
        F.cgR = cgR.p;
        F.cgP = cgP.p;
        F.cgZ = cgZ.p;
        F.cgAp = cgAp.p;
        F.partials = partials.p;
        F.partialsF = partialsF.p;
        F.live = live.p;
        F.nbrLocal = nbrLocal.p;
        F.haloCount = haloCount.p;
        F.haloIds = haloIds.p;
        F.state = dState.p;
        F.halton = dHalton.p;
        return F;
    }

    // ParticleSearch::ComputeMinMax + ComputeCellInformation (ParticleSearch.cu:29-93) on the host, per scene.
    int search() {
        const Launch L = launch();
        const uint32_t S = (uint32_t)sc.size();
        VFD_TRY(timed("minmax", [&] { launch_minmax(L, posNext.p, dMMPart.p, dMM.p); }));
        std::vector<int> mm((size_t)S * 6);
        VFD_HIP(hipMemcpyAsync(mm.data(), dMM.p, mm.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        uint32_t cellTotal = 0, prefix = 0;
        std::vector<SceneGrid> grids(S);
        for (uint32_t k = 0; k < S; ++k) {
            Scene& s = *sc[k];
            SceneGrid g{};
            const int32_t startCorr = (int32_t)s.base - (int32_t)prefix;
            prefix += s.count;
            if (s.count != 0) {
                const uint64_t cells = size_scene_grid(&mm[6 * k], &mm[6 * k + 3], s.prm.R, cellTotal, g);
                if (cells == 0)
                    return fail(VFD_ERROR_NOT_SUPPORTED,
                                "scene " + std::to_string(k) + ": particle cloud too sparse for the neighbour grid "
                                "(more than " + std::to_string(kMaxGridCells) + " cells)");
            }
            g.cellBase = cellTotal;
            g.startCorr = startCorr;
            cellTotal += g.cellCount;  // <= kMaxGridCells, checked by size_scene_grid
            grids[k] = g;
            s.grid = g;
        }
        uint32_t *cCount = nullptr, *cStart = nullptr, *cScan = nullptr;
        if (t_noAlloc) {  // the static scratch: checked before any kernel writes the particle state
            if (cellTotal > kStaticCells)
                return fail(VFD_ERROR_NOT_SUPPORTED,
                            "vfd_step_no_alloc: the neighbour grid needs " + std::to_string(cellTotal) +
                                " cells, more than the static scratch holds (" + std::to_string(kStaticCells) + ")");
            cCount = staticScratch + kStaticCellCountAt;
            cStart = staticScratch + kStaticCellStartAt;
            cScan = staticScratch + kStaticScanAt;
        } else {
            if (cellTotal > cellCap) {
                const uint32_t want = (uint32_t)grid_capacity(cellTotal);
                VFD_TRY(cellCount.alloc(want));
                VFD_TRY(cellStart.alloc(want));
                VFD_TRY(scanScratch.alloc(scan_scratch_words(want)));
                cellCap = want;
            }
            cCount = cellCount.p;
            cStart = cellStart.p;
            cScan = scanScratch.p;
        }
        VFD_HIP(hipMemcpyAsync(dGrid.p, grids.data(), S * sizeof(SceneGrid), hipMemcpyHostToDevice, stream));
        const int alt = cur ^ 1;
        VFD_TRY(timed("sort", [&] {
            launch_sort(L, dGrid.p, cellTotal, posNext.p, vel[cur].p, dv[cur].p, dk[cur].p, id[cur].p, pos.p,
                        vel[alt].p, dv[alt].p, dk[alt].p, id[alt].p, cellOf.p, rank.p, tmpId.p, cCount, cStart,
                        cScan);
        }));
        cur = alt;
        VFD_TRY(timed("build_lists", [&] {
            launch_build_lists(L, dGrid.p, pos.p, cStart, cCount, nbrCount.p, nbrList.p, nbStats.p);
        }));
        VFD_HIP(hipGetLastError());
        return VFD_OK;
    }

    // Runs a solver loop whose iterations are no-ops on the device once a scene has converged: `first` iterations,
    // then chunks of pollChunk. Each chunk ends with an asynchronous copy of live[which] into a pinned slot; the next
    // chunk is enqueued before the host waits for the previous slot, so the GPU never idles on a convergence check
    // (at most one chunk of empty launches after the last scene converges).
    template <class Fn>
    int poll_loop(Fn enqueue, uint32_t first, int which) {
        uint32_t iter = 0;
        // graphs are keyed by the double-buffer index of the particle arrays their launches captured
        const int buf = this->cur;
        auto issue = [&](uint32_t n, int slot) -> int {
            if (useGraphs && !profiling && !taping && n == pollChunk && stream != nullptr && !t_noAlloc) {
                if (graphsStale) drop_graphs();
                hipGraphExec_t& g = loopGraph[which][buf];
                if (!g) {
                    hipGraph_t graph = nullptr;
                    VFD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                    for (uint32_t c = 0; c < n; ++c) enqueue(iter + c);
                    VFD_HIP(hipStreamEndCapture(stream, &graph));
                    const hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    VFD_HIP(e);
                }
                VFD_HIP(hipGraphLaunch(g, stream));
                iter += n;
            } else {
                for (uint32_t c = 0; c < n; ++c) enqueue(iter++);
            }
            VFD_HIP(hipMemcpyAsync(hostLive + slot, live.p + which, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
            VFD_HIP(hipEventRecord(pollEv[slot], stream));
            return VFD_OK;
        };
        int slot = 0;  // pinned poll slot of the chunk in flight
        uint32_t n0 = std::max<uint32_t>(first, 1);
        if (predictIterations && lastIterations[which] > 1) n0 = std::max(n0, lastIterations[which] - 1);
        VFD_TRY(issue(n0, slot));
        for (;;) {
            VFD_TRY(issue(pollChunk, slot ^ 1));
            VFD_HIP(hipEventSynchronize(pollEv[slot]));
            if (hostLive[slot] == 0) return VFD_OK;
            slot ^= 1;
        }
    }

    // OnUpdate (DFSPHImplementation.cu:63-170), stopping after phase `last`.
    int step(int last) {
        if (!committed) VFD_TRY(commit());
        if (tiles == 0) return VFD_OK;  // OnUpdate returns for N = 0 (:65-67)
        const Launch L = launch();
        const KTabHost K = ktab();
        bool anyST = false, anyVisc = false;
        uint32_t minDiv = 0, minPres = 0, minVisc = 1;
        for (Scene* s : sc) {
            anyST = anyST || (s->desc.EnableSurfaceTensionSolver && s->count);
            anyVisc = anyVisc || (s->desc.EnableViscositySolver && s->count);
            minDiv = std::max(minDiv, s->desc.MinDivergenceSolverIterations);
            minPres = std::max(minPres, s->desc.MinPressureSolverIterations);
        }
        VFD_HIP(hipEventRecord(ev[0], stream));
        VFD_TRY(search());
        VFD_HIP(hipEventRecord(ev[1], stream));
        if (last == VFD_PHASE_SEARCH) return finish_partial();
        Fields F = fields();
        if (taping) VFD_TRY(tape_copy(tapeXpre, pos.p, cap));
        VFD_TRY(timed("boundary", [&] { launch_boundary(L, dMaps.p, pos.p, vel[cur].p, xb.p, vb.p); }));
        // halos are built after the boundary has pushed penetrating particles out (the lists themselves come from the
        // positions before)
        VFD_TRY(timed("build_halo", [&] {
            launch_build_halo(L, nbrCount.p, nbrList.p, haloLimit, haloCount.p, haloIds.p, nbrLocal.p, haloStats.p,
                              pos.p, K);
        }));
        if (last == VFD_PHASE_BOUNDARY) return finish_partial();
        DevTape* tp = nullptr;
        if (taping) {
            if (tapeCount == tapes.size()) tapes.emplace_back(new DevTape());
            tp = tapes[tapeCount++].get();
            const size_t n = cap;  // every scene's slots, each scene reversed over its own range
            tp->n = (uint32_t)n;
            tp->scn.assign(sc.size(), DevTape::SceneRec{});
            tp->st = tp->stTemporal = false;
            uint32_t maxBodies = 0;
            for (size_t s = 0; s < sc.size(); ++s) {
                DevTape::SceneRec& r = tp->scn[s];
                const vfd_desc& d = sc[s]->desc;
                r.h = sc[s]->st.dt;
                r.bodies = (uint32_t)sc[s]->bodies.size();
                r.boundaryViscosity = d.EnableViscositySolver && d.BoundaryViscosity != 0.0f;
                r.visc = d.EnableViscositySolver != 0;
                r.st = d.EnableSurfaceTensionSolver != 0;
                r.stPasses = d.SurfaceTensionSmoothPassCount;
                r.stTemporal = r.st && d.TemporalSmoothing != 0;
                tp->st = tp->st || r.st;
                tp->stTemporal = tp->stTemporal || r.stTemporal;
                maxBodies = std::max(maxBodies, r.bodies);
            }
            VFD_TRY(tape_copy(tp->x, pos.p, n));
            VFD_TRY(tape_copy(tp->v0, vel[cur].p, n));
            VFD_TRY(tape_copy(tp->id, id[cur].p, n));
            VFD_TRY(tape_copy(tp->count, nbrCount.p, n));
            VFD_TRY(tape_copy(tp->list, nbrList.p, (size_t)cap * kMaxNeighbors));
            // boundary pass replayed with its derivatives; scope checked by vfd_adjoint from the flags
            VFD_TRY(tp->flags.alloc(3 * sc.size()));
            VFD_HIP(hipMemsetAsync(tp->flags.p, 0, 3 * sc.size() * sizeof(uint32_t), stream));
            if (maxBodies) {
                VFD_TRY(tp->con.alloc((size_t)kContactRecs * kMaxBodies * n));
                VFD_TRY(tp->push.alloc((size_t)3 * n));
                adj_launch_boundary_tape(stream, dP.p, dTileScene.p, dMaps.p, (uint32_t)n, cap, tapeXpre.p, pos.p,
                                         vb.p, xb.p, tp->con.p, tp->push.p, tp->flags.p);
            }
        }
        VFD_TRY(timed("density_factor", [&] { launch_density_factor(L, K, F); }));
        if (tp) {
            VFD_TRY(tape_copy(tp->rho, density.p, tp->n));
            VFD_TRY(tape_copy(tp->alpha, factor.p, tp->n));
        }
        VFD_HIP(hipEventRecord(ev[2], stream));
        if (last == VFD_PHASE_DENSITY || last == VFD_PHASE_FACTOR) return finish_partial();
        // divergence solve (ComputeDivergence :506-575)
        launch_scene_init_div(L, F);
        VFD_TRY(timed("div_init", [&] { launch_div_init(L, K, F); }));
        if (tp) VFD_TRY(tape_copy(tp->D, densAdv.p, tp->n));
        VFD_TRY(poll_loop(
            [&](uint32_t k) {
                if (tp && k == 0) (void)tape_slot(tp->kv, 0, kappaV.p, tp->n);
                (void)timed("div_accel", [&] { launch_div_accel(L, K, F); }, 0, k);
                (void)timed("div_iter", [&] { launch_div_iter(L, K, F); }, 0, k);
                (void)timed("jacobi_finish", [&] { launch_jacobi_finish(L, F, 0); }, 0, k);
                if (tp) (void)tape_iteration(tp->kv, tp->av, tp->rv, k, tp->n, kappaV.p, ap.p, residuum.p);
            },
            minDiv, 0));
        if (tp) VFD_TRY(tape_copy(tp->kvEnd, kappaV.p, tp->n));
        VFD_TRY(timed("div_final", [&] { launch_div_final(L, K, F); }));  // also ClearAccelerationKernel
        if (tp) VFD_TRY(tape_copy(tp->v1, vel[cur].p, tp->n));
        VFD_HIP(hipEventRecord(ev[3], stream));
        if (last == VFD_PHASE_DIVERGENCE || last == VFD_PHASE_CLEAR_ACCELERATION) return finish_partial();
        if (anyST) {
            if (tp && tp->stTemporal) VFD_TRY(tape_copy(tp->dK0, dk[cur].p, tp->n));
            VFD_TRY(upload_mc_tables(F));
            VFD_TRY(timed("surface_tension", [&] { launch_surface_tension(L, F); }));
        }
        if (tp && tp->st) {
            VFD_TRY(tape_copy(tp->nMC, nMC.p, tp->n));
            VFD_TRY(tape_copy(tp->kMC, kMC.p, tp->n));
        }
        VFD_HIP(hipEventRecord(ev[4], stream));
        if (last == VFD_PHASE_SURFACE_TENSION) return finish_partial();
        if (anyVisc) {
            VFD_TRY(timed("visc_precond", [&] { launch_visc_precond(L, K, F); }));
            VFD_TRY(timed("cg_begin", [&] { launch_cg_begin(L, K, F); }));
            VFD_TRY(timed("cg_begin_finish", [&] { launch_cg_begin_finish(L, F); }));
            VFD_TRY(poll_loop(
                [&](uint32_t k) {
                    (void)timed("cg_matvec", [&] { launch_cg_matvec(L, K, F); }, 2, k);
                    (void)timed("cg_alpha", [&] { launch_cg_alpha(L, F); }, 2, k);
                    (void)timed("cg_update", [&] { launch_cg_update(L, F); }, 2, k);
                    (void)timed("cg_update_finish", [&] { launch_cg_update_finish(L, F); }, 2, k);
                    (void)timed("cg_direction", [&] { launch_cg_direction(L, F); }, 2, k);
                },
                minVisc, 2));
            VFD_TRY(timed("visc_apply", [&] { launch_visc_apply(L, F); }));
            if (tp) VFD_TRY(tape_copy(tp->y, cgX.p, tp->n));
        }
        VFD_HIP(hipEventRecord(ev[5], stream));
        if (last == VFD_PHASE_VISCOSITY) return finish_partial();
        VFD_TRY(timed("time_step", [&] { launch_time_step(L, F); }));
        VFD_TRY(timed("time_step_finish", [&] { launch_time_step_finish(L, F); }));
        if (last == VFD_PHASE_TIME_STEP) return finish_partial();
        VFD_TRY(timed("velocity", [&] { launch_velocity(L, F); }));
        if (tp) VFD_TRY(tape_copy(tp->v2, vel[cur].p, tp->n));
        if (last == VFD_PHASE_VELOCITY) return finish_partial();
        launch_scene_init_pres(L, F);
        VFD_TRY(timed("pres_init", [&] { launch_pres_init(L, K, F); }));
        if (tp) VFD_TRY(tape_copy(tp->rhoadv, densAdv.p, tp->n));
        VFD_TRY(poll_loop(
            [&](uint32_t k) {
                if (tp && k == 0) (void)tape_slot(tp->kp, 0, kappa.p, tp->n);
                (void)timed("pres_accel", [&] { launch_pres_accel(L, K, F); }, 1, k);
                (void)timed("pres_iter", [&] { launch_pres_iter(L, K, F); }, 1, k);
                (void)timed("jacobi_finish", [&] { launch_jacobi_finish(L, F, 1); }, 1, k);
                if (tp) (void)tape_iteration(tp->kp, tp->ap, tp->rp, k, tp->n, kappa.p, ap.p, residuum.p);
            },
            minPres, 1));
        if (tp) VFD_TRY(tape_copy(tp->kpEnd, kappa.p, tp->n));
        VFD_TRY(timed("pres_final", [&] { launch_pres_final(L, K, F, last >= VFD_PHASE_POSITION ? 1 : 0); }));
        VFD_HIP(hipEventRecord(ev[6], stream));
        if (last == VFD_PHASE_PRESSURE) return finish_partial();
        VFD_HIP(hipGetLastError());
        VFD_HIP(hipStreamSynchronize(stream));
        std::unique_lock<std::mutex> lk(hostMu);  // the host bookkeeping that polling threads read
        VFD_TRY(pull_state());
        solver_debug();
        VFD_TRY(neighbor_stats());
        VFD_TRY(resolve_profile());
        int nonfiniteScene = -1;  // per-step non-finite check (no reference counterpart)
        for (size_t k = 0; k < sc.size() && nonfiniteScene < 0; ++k)
            if (sc[k]->count && sc[k]->st.nonfinite) nonfiniteScene = (int)k;
        if (tp)
            for (size_t s = 0; s < sc.size(); ++s) {
                DevTape::SceneRec& r = tp->scn[s];
                r.H = sc[s]->st.dt;
                r.Kd = sc[s]->desc.EnableDivergenceSolverError ? sc[s]->st.div.it : 0;
                r.Kp = sc[s]->st.pres.it;
                r.atCap = sc[s]->count && sc[s]->dbg.MaxNeighborCount >= kMaxNeighbors;
            }
        float t[6] = {0, 0, 0, 0, 0, 0};
        VFD_HIP(hipEventElapsedTime(&t[0], ev[0], ev[1]));
        VFD_HIP(hipEventElapsedTime(&t[1], ev[1], ev[2]));
        VFD_HIP(hipEventElapsedTime(&t[2], ev[2], ev[3]));
        VFD_HIP(hipEventElapsedTime(&t[3], ev[3], ev[4]));
        VFD_HIP(hipEventElapsedTime(&t[4], ev[4], ev[5]));
        VFD_HIP(hipEventElapsedTime(&t[5], ev[5], ev[6]));
        for (Scene* s : sc) {
            if (s->count == 0) continue;
            vfd_debug_info& d = s->dbg;
            d.NeighborhoodSearchTime = t[0];
            d.BaseSolverTime = t[1];
            d.DivergenceSolverTime = t[2];
            d.SurfaceTensionSolverTime = t[3];
            d.ViscositySolverTime = t[4];
            d.PressureSolverTime = t[5];
            d.FrameTime += s->st.dt;  // ComputeTimeStepSize :427
        }
        // frame capture (:148-167)
        for (size_t k = 0; k < sc.size(); ++k) {
            Scene* s = sc[k];
            if (s->count == 0) continue;
            vfd_debug_info& d = s->dbg;
            if (d.FrameIndex < s->desc.FrameCount && d.FrameTime >= s->desc.FrameLength) {
                if (capture) {
                    Frame fr;
                    fr.maxVel = s->st.maxVel;
                    fr.dt = s->st.dt;
                    const int rc = capture_frame((uint32_t)k, fr);
                    s->frames.push_back(std::move(fr));  // (owned by the scene even on failure, so it is released)
                    VFD_TRY(rc);
                }
                d.FrameTime = 0.0f;
                d.FrameIndex++;
            }
            d.IterationCount++;
        }
        if (nonfiniteScene >= 0)
            return fail(VFD_ERROR_NONFINITE, "scene " + std::to_string(nonfiniteScene) +
                                                 ": a particle's position or velocity is not finite after step " +
                                                 std::to_string(sc[nonfiniteScene]->dbg.IterationCount));
        return VFD_OK;
    }

    // ---- adjoint: reverse of the recorded steps (kernels in adjoint.hip) ----------------------------------------
    DevBuf<float4> aLx, aLv, aLa, aTmp4, aGs, aLy, aLb, bR, bRh, bP, bV, bS, bT, aLnc;
    DevBuf<float2> aLcw;
    DevBuf<float> aLk, aLsrc, aLs, aLf, aLrho, aLalpha, aLd, aLS, aOrigX, aOrigV, aOrigK, aLdK;
    DevBuf<double> aPart, aMu;  // aMu: d/dmu accumulated over the reversed steps
    // halos of the step being reversed, rebuilt from its recorded lists for the group kernels
    DevBuf<uint32_t> aHaloCount, aHaloIds, aNbrLocal, aHaloStats;
    // transposed Jacobi and viscosity loops: 0 per particle, 1 per group, 2 per group once there is a group for every
    // CU (below that the per-particle kernels spread better and run as fast: 100k particles, 98 groups)
    uint32_t adjointGroups = 2;
    // BiCGSTAB for A^T x = b (the transposed viscosity system of one recorded step), relative residual 1e-6. The
    // iterations run on the device with their scalars in a BicgState; the host reads the state every
    // kBicgCheck iterations.
    DevBuf<BicgState> dBicg;
    BicgState hBicg{};
    static constexpr int kBicgCheck = 4;
    int adj_visc_solve(const AdjKHost& A, float h, const float* rho, const float4* b, float4* x) {
        // slot-indexed vectors (A.stride slots); the scene's rows are [A.base, A.base + A.n)
        const uint32_t n = A.n, o = A.base;
        for (auto* q : {&bR, &bRh, &bP, &bV, &bS, &bT}) VFD_TRY(q->alloc(A.stride));
        VFD_TRY(dBicg.alloc(1));
        VFD_HIP(hipMemcpyAsync(x + o, b + o, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
        adj_launch_visc_T(stream, A, h, rho, x, bT.p);
        VFD_HIP(hipMemcpyAsync(bR.p + o, b + o, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
        adj_launch_axpy4(stream, n, -1.0f, bT.p + o, bR.p + o);
        VFD_HIP(hipMemcpyAsync(bRh.p + o, bR.p + o, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
        VFD_HIP(hipMemsetAsync(bP.p + o, 0, n * sizeof(float4), stream));
        VFD_HIP(hipMemsetAsync(bV.p + o, 0, n * sizeof(float4), stream));
        // |b|^2, |r0|^2 and the threshold on the device (done at once when b = 0 or r0 = b - A^T b meets it)
        VFD_TRY(aPart.alloc(2 * (size_t)((n + kTile - 1) / kTile)));
        adj_launch_bicg_begin(stream, A, b, bR.p, aPart.p, dBicg.p, 300u);
        const BicgVectors vec{x, bR.p, bRh.p, bP.p, bV.p, bS.p, bT.p};
        for (uint32_t done = 0; !done;) {
            for (int k = 0; k < kBicgCheck; ++k) adj_launch_bicg_iteration(stream, A, h, rho, vec, aPart.p, dBicg.p);
            VFD_HIP(hipGetLastError());
            VFD_HIP(hipMemcpyAsync(&hBicg, dBicg.p, sizeof(BicgState), hipMemcpyDeviceToHost, stream));
            VFD_HIP(hipStreamSynchronize(stream));
            done = hBicg.done || hBicg.it >= hBicg.maxIt;
        }
        return VFD_OK;
    }
    // reverse of scene sc_ of one recorded step: aLx, aLv hold the cotangents of its output (x', v') on entry and
    // those of its input (x, v) on return, in the step's slot order (the scene's particles in slots
    // [base, base + count)), and so does ldK (null: not tracked) for the blended curvature the surface tension's
    // temporal smoothing carries; lmu (one device double) accumulates d/dmu. Neighbour kernels take slot-indexed
    // arrays whole (the lists hold slots), pointwise kernels take them from the scene's first slot.
    int adj_step(const DevTape& t, uint32_t sc_, const uint32_t flags[3], double* lmu, float* ldK) {
        const Scene& S = *sc[sc_];
        const DevTape::SceneRec& r = t.scn[sc_];
        const uint32_t N = t.n, n = S.count, o = S.base;
        const SceneParams& pr = S.prm;
        AdjKHost A{dTab.p, kt.R, kt.invStep, kt.K, kt.L, pr.V, pr.rho0, pr.m, pr.mu, kt.R2, n, cap,
                   t.count.p, t.list.p, t.x.p};
        A.base = o;
        A.stride = N;
        if (flags[0]) {  // some particle sampled a body: the sample terms join every operator
            A.con = t.con.p;
            A.bodies = r.bodies;
            A.muB = r.boundaryViscosity ? pr.muB : 0.0f;
            A.td = pr.td;
        }
        const Launch L = launch();
        if (sc.size() == 1 && (adjointGroups == 1 || (adjointGroups == 2 && L.groups >= numCU))) {
            VFD_TRY(aHaloCount.alloc(std::max<uint32_t>(L.groups, 1)));
            VFD_TRY(aHaloIds.alloc((size_t)std::max<uint32_t>(L.groups, 1) * kHaloMax));
            VFD_TRY(aNbrLocal.alloc((size_t)cap * kMaxNeighbors));
            VFD_TRY(aHaloStats.alloc(3));
            launch_build_halo(L, t.count.p, t.list.p, haloLimit, aHaloCount.p, aHaloIds.p, aNbrLocal.p, aHaloStats.p,
                              t.x.p, ktab());
            A.groups.grid = L.gridGroups;
            A.groups.tiles = L.tg.tiles;
            A.groups.per = L.tg.per;
            A.groups.haloCount = aHaloCount.p;
            A.groups.haloIds = aHaloIds.p;
            A.groups.local = aNbrLocal.p;
        }
        const float h = r.h, H = r.H;
        for (auto* q : {&aLa, &aTmp4, &aGs, &aLy, &aLb}) VFD_TRY(q->alloc(N));
        for (auto* q : {&aLk, &aLsrc, &aLs, &aLf, &aLrho, &aLalpha, &aLd, &aLS}) VFD_TRY(q->alloc(N));
        float4 *lx = aLx.p + o, *lv = aLv.p + o, *la = aLa.p + o, *ly = aLy.p + o, *lb = aLb.p + o;
        float *lk = aLk.p + o, *lsrc = aLsrc.p + o, *ls = aLs.p + o, *lf = aLf.p + o, *lrho = aLrho.p + o;
        float *lalpha = aLalpha.p + o, *ld = aLd.p + o;
        adj_launch_fill(stream, lrho, 0.0f, n);
        adj_launch_fill(stream, lalpha, 0.0f, n);
        // x' = x + H v3  ->  lv3 = lv + H lx (aLv becomes lv3, then lv2)
        adj_launch_axpy4(stream, n, H, lx, lv);
        // pressure final: v3 = v2 + H A(kpEnd)
        adj_launch_scale4(stream, n, lv, H, la);
        adj_launch_fill(stream, lk, 0.0f, n);
        adj_launch_accel_T(stream, A, t.kpEnd.p, aLa.p, aLk.p, aLx.p);
        // pressure iterations, gamma = alpha / H^2, source 1 - rhoadv, c = H^2
        adj_launch_fill(stream, lsrc, 0.0f, n);
        adj_launch_fill(stream, ls, 0.0f, n);
        for (int k = (int)r.Kp - 1; k >= 0; --k) {
            adj_launch_jacobi_point(stream, n, lk, t.alpha.p + o, H * H, t.rp[k].p + o, H * H, lsrc, ls, lf);
            VFD_HIP(hipMemsetAsync(la, 0, n * sizeof(float4), stream));
            adj_launch_div_T(stream, A, aLf.p, t.ap[k].p, aLa.p, aLx.p);
            adj_launch_accel_T(stream, A, t.kp[k].p, aLa.p, aLk.p, aLx.p);
        }
        adj_launch_pres_init(stream, n, H, pr.rho0, t.alpha.p + o, t.rhoadv.p + o, lk, lsrc, ls, lalpha, lrho, ld);
        adj_launch_div_T(stream, A, aLd.p, t.v2.p, aLv.p, aLx.p);  // rhoadv's H D(v2)
        // v2 includes H a_st(x): surface tension with its Monte-Carlo normals and curvatures held
        if (r.st && r.stPasses) {
            VFD_TRY(aLnc.alloc(N));
            VFD_TRY(aLcw.alloc(N));
            adj_launch_surface_tension(stream, A, H, aLv.p, t.nMC.p, t.kMC.p, pr.smoothing, -pr.sigma * pr.mInv,
                                       r.stPasses, r.stTemporal ? t.dK0.p : nullptr, ldK, aLnc.p, aLcw.p, aLx.p);
        }
        // v2 = v1 + H g + (H / h)(y - v1)
        if (r.visc) {
            adj_launch_scale4(stream, n, lv, H / h, ly);
            adj_launch_scale4(stream, n, lv, 1.0f - H / h, lv);
            // A y = v1: lb = A^-T ly; lv1 += lb; positions, densities and mu through lb . dT
            VFD_TRY(adj_visc_solve(A, h, t.rho.p, aLy.p, aLb.p));
            adj_launch_axpy4(stream, n, 1.0f, lb, lv);
            const uint32_t blocks = (n + kTile - 1) / kTile;
            VFD_TRY(aPart.alloc(blocks));
            adj_launch_visc_vjp(stream, A, h, t.rho.p, t.y.p, aLb.p, aLx.p, aLrho.p, aPart.p);
            adj_launch_mu_acc(stream, aPart.p, blocks, lmu);
        }
        // divergence final: v1 = v0 + h A(kvEnd)
        adj_launch_scale4(stream, n, lv, h, la);
        adj_launch_fill(stream, lk, 0.0f, n);
        adj_launch_accel_T(stream, A, t.kvEnd.p, aLa.p, aLk.p, aLx.p);
        // divergence iterations, beta = alpha / h, source -D, c = h
        adj_launch_fill(stream, lsrc, 0.0f, n);
        adj_launch_fill(stream, ls, 0.0f, n);
        for (int k = (int)r.Kd - 1; k >= 0; --k) {
            adj_launch_jacobi_point(stream, n, lk, t.alpha.p + o, h, t.rv[k].p + o, h, lsrc, ls, lf);
            VFD_HIP(hipMemsetAsync(la, 0, n * sizeof(float4), stream));
            adj_launch_div_T(stream, A, aLf.p, t.av[k].p, aLa.p, aLx.p);
            adj_launch_accel_T(stream, A, t.kv[k].p, aLa.p, aLk.p, aLx.p);
        }
        adj_launch_div_init(stream, n, h, t.alpha.p + o, t.D.p + o, lk, lsrc, ls, lalpha, ld);
        adj_launch_div_T(stream, A, aLd.p, t.v0.p, aLv.p, aLx.p);
        // factor and density
        adj_launch_factor_density(stream, A, t.alpha.p, aLalpha.p, aLS.p, aGs.p, aLrho.p, aLx.p);
        // boundary push-out (the step's positions and velocities above are those after it)
        if (flags[1]) adj_launch_push(stream, n, N, t.push.p + o, lx, lv);
        VFD_HIP(hipGetLastError());
        return VFD_OK;
    }
    // cotangents of scene sc_'s particles (original order) after the last recorded step -> before the first
    int adjoint(uint32_t sc_, const float* lxOut, const float* lvOut, float* lxIn, float* lvIn, double* lmu) {
        if (tapeCount == 0) return fail(VFD_ERROR_STATE, "no recorded steps (vfd_tape)");
        if (sc_ >= sc.size()) return fail(VFD_ERROR_INVALID_ARGUMENT, "scene index out of range");
        for (size_t k = 0; k < tapeCount; ++k)
            if (tapes[k]->scn.size() != sc.size())
                return fail(VFD_ERROR_STATE, "the scenes changed since the steps were recorded");
        std::vector<uint32_t> flags(3 * tapeCount);
        for (size_t k = 0; k < tapeCount; ++k)
            VFD_HIP(hipMemcpyAsync(&flags[3 * k], tapes[k]->flags.p + 3 * (size_t)sc_, 3 * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        bool carry = false;  // some step's temporal smoothing reads the curvature carried from the step before
        std::string err = tapeError;
        for (size_t k = 0; k < tapeCount; ++k) {
            const DevTape::SceneRec& r = tapes[k]->scn[sc_];
            carry = carry || r.stTemporal;
            if (flags[3 * k + 2])
                err = "a particle pushed out twice or after a contact sample is outside the adjoint's scope";
            if (r.atCap) err = "a particle reached the neighbour cap (lists not symmetric)";
        }
        if (!err.empty()) return fail(VFD_ERROR_NOT_SUPPORTED, err);
        const Scene& S = *sc[sc_];
        const uint32_t N = tapes[tapeCount - 1]->n, n = S.count, o = S.base;
        if (n == 0) return fail(VFD_ERROR_STATE, "the scene has no particles");
        VFD_TRY(aOrigX.alloc((size_t)3 * n));
        VFD_TRY(aOrigV.alloc((size_t)3 * n));
        VFD_TRY(aLx.alloc(N));
        VFD_TRY(aLv.alloc(N));
        VFD_HIP(hipMemcpyAsync(aOrigX.p, lxOut, (size_t)3 * n * sizeof(float), hipMemcpyHostToDevice, stream));
        if (lvOut)
            VFD_HIP(hipMemcpyAsync(aOrigV.p, lvOut, (size_t)3 * n * sizeof(float), hipMemcpyHostToDevice, stream));
        else
            VFD_HIP(hipMemsetAsync(aOrigV.p, 0, (size_t)3 * n * sizeof(float), stream));
        if (carry) {  // the loss reads x and v: the carried curvature after the last step has cotangent 0
            VFD_TRY(aOrigK.alloc(n));
            VFD_TRY(aLdK.alloc(N));
            adj_launch_fill(stream, aOrigK.p, 0.0f, n);
        }
        VFD_TRY(aMu.alloc(1));
        VFD_HIP(hipMemsetAsync(aMu.p, 0, sizeof(double), stream));
        for (size_t k = tapeCount; k-- > 0;) {
            const DevTape& t = *tapes[k];
            adj_launch_to_slots(stream, n, t.id.p + o, aOrigX.p, aLx.p + o);
            adj_launch_to_slots(stream, n, t.id.p + o, aOrigV.p, aLv.p + o);
            if (carry) adj_launch_to_slots1(stream, n, t.id.p + o, aOrigK.p, aLdK.p + o);
            VFD_TRY(adj_step(t, sc_, &flags[3 * k], aMu.p, carry ? aLdK.p : nullptr));
            adj_launch_to_orig(stream, n, t.id.p + o, aLx.p + o, aOrigX.p);
            adj_launch_to_orig(stream, n, t.id.p + o, aLv.p + o, aOrigV.p);
            if (carry) adj_launch_to_orig1(stream, n, t.id.p + o, aLdK.p + o, aOrigK.p);
        }
        VFD_HIP(hipMemcpyAsync(lxIn, aOrigX.p, (size_t)3 * n * sizeof(float), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipMemcpyAsync(lvIn, aOrigV.p, (size_t)3 * n * sizeof(float), hipMemcpyDeviceToHost, stream));
        double m = 0;
        VFD_HIP(hipMemcpyAsync(&m, aMu.p, sizeof(double), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        if (lmu) *lmu = m;
        return VFD_OK;
    }

    // after a partial step the working positions become the state the next step starts from
    int finish_partial() {
        VFD_HIP(hipMemcpyAsync(posNext.p, pos.p, (size_t)cap * sizeof(float4), hipMemcpyDeviceToDevice, stream));
        VFD_HIP(hipGetLastError());
        VFD_HIP(hipStreamSynchronize(stream));
        std::lock_guard<std::mutex> lk(hostMu);
        VFD_TRY(pull_state());
        solver_debug();
        VFD_TRY(neighbor_stats());
        return resolve_profile();
    }

    int neighbor_stats() {
        std::vector<unsigned long long> nb(2 * sc.size());
        VFD_HIP(hipMemcpyAsync(nb.data(), nbStats.p, nb.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               stream));
        VFD_HIP(hipStreamSynchronize(stream));
        uint32_t hs[3];
        VFD_HIP(hipMemcpyAsync(hs, haloStats.p, sizeof(hs), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        for (size_t k = 0; k < sc.size(); ++k) {
            sc[k]->dbg.TotalNeighborCount = nb[2 * k];
            sc[k]->dbg.MaxNeighborCount = (uint32_t)nb[2 * k + 1];
            sc[k]->dbg.StagedGroupCount = hs[0];
            sc[k]->dbg.GatheredGroupCount = hs[1];
            sc[k]->dbg.MaxHaloSize = hs[2];
        }
        return VFD_OK;
    }

    // solver iteration counts and errors of the last (possibly partial) step, DFSPHDebugInfo semantics
    void solver_debug() {
        uint32_t ran[3] = {0, 0, 0};
        for (Scene* s : sc) {
            if (!s->count) continue;
            ran[0] = std::max(ran[0], s->st.div.it);
            ran[1] = std::max(ran[1], s->st.pres.it);
            ran[2] = std::max(ran[2], s->st.cg.loops);
        }
        std::memcpy(lastIterations, ran, sizeof(ran));
        for (Scene* s : sc) {
            vfd_debug_info& d = s->dbg;
            d.DivergenceSolverIterationCount = s->desc.EnableDivergenceSolverError ? s->st.div.it : 0;
            d.DivergenceSolverError = s->desc.EnableDivergenceSolverError ? s->st.div.err : 0.0f;
            d.PressureSolverIterationCount = s->st.pres.it;
            d.PressureSolverError = s->st.pres.err;
            d.ViscositySolverIterationCount = s->desc.EnableViscositySolver ? s->st.cg.it : 0;
            d.ViscositySolverError = s->desc.EnableViscositySolver ? s->st.cg.err : 0.0f;
            d.TimeStepSize = s->st.dt;
            d.MaxVelocityMagnitude = s->st.maxVel;
        }
    }

    // ---- reads in original order --------------------------------------------------------------------------------
    template <class T> int fetch(const T* dptr, uint32_t s, std::vector<T>& out) {
        const Scene& S = *sc[s];
        out.resize(S.count);
        if (S.count)
            VFD_HIP(hipMemcpyAsync(out.data(), dptr + S.base, S.count * sizeof(T), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        return VFD_OK;
    }
    int ids(uint32_t s, std::vector<uint32_t>& out) { return fetch(id[cur].p, s, out); }

    int get_field(uint32_t s, int field, uint32_t body, float* out) {
        if (!committed) VFD_TRY(commit());
        const Scene& S = *sc[s];
        std::vector<uint32_t> idv;
        VFD_TRY(ids(s, idv));
        auto put4 = [&](const float4* d, int comps) -> int {
            std::vector<float4> v;
            VFD_TRY(fetch(d, s, v));
            for (uint32_t k = 0; k < S.count; ++k) {
                const uint32_t i = idv[k];
                out[comps * i] = v[k].x;
                if (comps > 1) {
                    out[comps * i + 1] = v[k].y;
                    out[comps * i + 2] = v[k].z;
                }
            }
            return VFD_OK;
        };
        auto put1 = [&](const float* d) -> int {
            std::vector<float> v;
            VFD_TRY(fetch(d, s, v));
            for (uint32_t k = 0; k < S.count; ++k) out[idv[k]] = v[k];
            return VFD_OK;
        };
        switch (field) {
            case VFD_FIELD_POSITION: return put4(posNext.p, 3);
            case VFD_FIELD_VELOCITY: return put4(vel[cur].p, 3);
            case VFD_FIELD_ACCELERATION: return put4(acc.p, 3);
            case VFD_FIELD_PRESSURE_ACCELERATION: return put4(ap.p, 3);
            case VFD_FIELD_PRESSURE_RESIDUUM: return put1(residuum.p);
            case VFD_FIELD_DENSITY: return put1(density.p);
            case VFD_FIELD_DENSITY_ADVECTION: return put1(densAdv.p);
            case VFD_FIELD_PRESSURE_RHO2: return put1(kappa.p);
            case VFD_FIELD_PRESSURE_RHO2_V: return put1(kappaV.p);
            case VFD_FIELD_FACTOR: return put1(factor.p);
            case VFD_FIELD_VELOCITY_DIFFERENCE: return put4(dv[cur].p, 3);
            case VFD_FIELD_MC_SURFACE_NORMAL: return put4(nMC.p, 3);
            case VFD_FIELD_MC_SURFACE_NORMAL_SMOOTH: return put4(nS.p, 3);
            case VFD_FIELD_MC_SURFACE_CURVATURE: return put1(kMC.p);
            case VFD_FIELD_MC_SURFACE_CURVATURE_SMOOTH: return put1(kS.p);
            case VFD_FIELD_DELTA_FINAL_CURVATURE: return put1(dk[cur].p);
            case VFD_FIELD_BOUNDARY_XJ:
                if (body >= kMaxBodies) return fail(VFD_ERROR_INVALID_ARGUMENT, "body out of range");
                return put4(xb.p + (size_t)body * cap, 3);
            case VFD_FIELD_BOUNDARY_VOLUME:
                if (body >= kMaxBodies) return fail(VFD_ERROR_INVALID_ARGUMENT, "body out of range");
                return put1(vb.p + (size_t)body * cap);
            case VFD_FIELD_VISCOSITY_SOLUTION: return put4(cgX.p, 3);
            case VFD_FIELD_VISCOSITY_PRECONDITIONER: {
                for (int e = 0; e < 9; ++e) {
                    std::vector<float> v;
                    VFD_TRY(fetch(minv.p + (size_t)e * cap, s, v));
                    for (uint32_t k = 0; k < S.count; ++k) out[9 * idv[k] + e] = v[k];
                }
                return VFD_OK;
            }
        }
        return fail(VFD_ERROR_INVALID_ARGUMENT, "unknown field");
    }

    int set_field(uint32_t s, int field, const float* in) {
        if (!committed) VFD_TRY(commit());
        const Scene& S = *sc[s];
        std::vector<uint32_t> idv;
        VFD_TRY(ids(s, idv));
        auto write4 = [&](float4* d) -> int {
            std::vector<float4> v(S.count);
            for (uint32_t k = 0; k < S.count; ++k) {
                const uint32_t i = idv[k];
                v[k] = make_float4(in[3 * i], in[3 * i + 1], in[3 * i + 2], 0.0f);
            }
            if (S.count) VFD_HIP(hipMemcpy(d + S.base, v.data(), S.count * sizeof(float4), hipMemcpyHostToDevice));
            return VFD_OK;
        };
        switch (field) {
            case VFD_FIELD_POSITION:
                VFD_TRY(write4(posNext.p));
                return write4(pos.p);
            case VFD_FIELD_VELOCITY: return write4(vel[cur].p);
            case VFD_FIELD_VELOCITY_DIFFERENCE: return write4(dv[cur].p);
            case VFD_FIELD_DELTA_FINAL_CURVATURE: {
                std::vector<float> v(S.count);
                for (uint32_t k = 0; k < S.count; ++k) v[k] = in[idv[k]];
                if (S.count)
                    VFD_HIP(hipMemcpy(dk[cur].p + S.base, v.data(), S.count * sizeof(float), hipMemcpyHostToDevice));
                return VFD_OK;
            }
        }
        return fail(VFD_ERROR_INVALID_ARGUMENT, "field is not writable");
    }

    int get_neighbors(uint32_t s, uint32_t* counts, uint32_t* offsets, uint32_t* out, uint64_t* total) {
        if (!committed) VFD_TRY(commit());
        const Scene& S = *sc[s];
        std::vector<uint32_t> idv, cnt;
        VFD_TRY(ids(s, idv));
        VFD_TRY(fetch(nbrCount.p, s, cnt));
        std::vector<uint32_t> cOrig(S.count), off(S.count);
        for (uint32_t k = 0; k < S.count; ++k) cOrig[idv[k]] = cnt[k];
        uint64_t run = 0;
        for (uint32_t i = 0; i < S.count; ++i) {
            off[i] = (uint32_t)run;
            run += cOrig[i];
        }
        if (total) *total = run;
        if (counts) std::memcpy(counts, cOrig.data(), S.count * sizeof(uint32_t));
        if (offsets) std::memcpy(offsets, off.data(), S.count * sizeof(uint32_t));
        if (!out) return VFD_OK;
        std::vector<uint32_t> rows((size_t)kMaxNeighbors * S.count);
        for (uint32_t j = 0; j < kMaxNeighbors; ++j)
            if (S.count)
                VFD_HIP(hipMemcpyAsync(&rows[(size_t)j * S.count], nbrList.p + (size_t)j * cap + S.base,
                                       S.count * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        VFD_HIP(hipStreamSynchronize(stream));
        for (uint32_t k = 0; k < S.count; ++k) {
            const uint32_t i = idv[k];
            for (uint32_t j = 0; j < cnt[k]; ++j) {
                const uint32_t t = rows[(size_t)j * S.count + k];
                if (t < S.base || t >= S.base + S.count) return fail(VFD_ERROR_STATE, "neighbour outside its scene");
                out[off[i] + j] = idv[t - S.base];
            }
        }
        return VFD_OK;
    }
};

}  // namespace vfd

using vfd::Simulation;
using vfd::fail;

struct vfd_sim {
    Simulation impl;
};

#define CHECK_SIM(sim)                                                                 \
    do {                                                                               \
        if (!(sim)) return fail(VFD_ERROR_INVALID_ARGUMENT, "vfd_sim handle is NULL"); \
    } while (0)
#define CHECK_SCENE(sim, s)                                                                   \
    do {                                                                                      \
        CHECK_SIM(sim);                                                                       \
        if ((s) >= (sim)->impl.sc.size()) return fail(VFD_ERROR_INVALID_ARGUMENT, "scene index out of range"); \
    } while (0)

static int guard(int rc) { return rc; }

// Held by every driving call for its duration: a call from a second thread while one runs (or while the worker of
// vfd_simulate_async runs) fails with VFD_ERROR_STATE.
struct DriveGuard {
    std::atomic<bool>& flag;
    bool ok;
    explicit DriveGuard(std::atomic<bool>& f) : flag(f) {
        bool idle = false;
        ok = flag.compare_exchange_strong(idle, true);
    }
    ~DriveGuard() {
        if (ok) flag.store(false);
    }
};
#define DRIVE(sim)                                                                                             \
    CHECK_SIM(sim);                                                                                            \
    DriveGuard drive_((sim)->impl.busy);                                                                       \
    if (!drive_.ok)                                                                                            \
        return fail(VFD_ERROR_STATE, "the simulation is running on another thread (vfd_simulate, vfd_step or "  \
                                     "vfd_simulate_async without vfd_wait)")
#define DRIVE_SCENE(sim, s)                                                                                   \
    DRIVE(sim);                                                                                               \
    if ((s) >= (sim)->impl.sc.size()) return fail(VFD_ERROR_INVALID_ARGUMENT, "scene index out of range")

extern "C" {

int vfd_abi_version(void) { return VFD_ABI_VERSION; }
const char* vfd_last_error(void) { return vfd::g_lastError.c_str(); }

int vfd_desc_default(vfd_desc* d) {
    if (!d) return fail(VFD_ERROR_INVALID_ARGUMENT, "desc is NULL");
    *d = vfd_desc{};
    d->TimeStepSize = 0.001f;
    d->MinTimeStepSize = 0.0001f;
    d->MaxTimeStepSize = 0.005f;
    d->FrameLength = 0.0016f;
    d->FrameCount = 200u;
    d->MinPressureSolverIterations = 0u;
    d->MaxPressureSolverIterations = 100u;
    d->MaxPressureSolverError = 10.0f;
    d->EnableDivergenceSolverError = 1;
    d->MinDivergenceSolverIterations = 0u;
    d->MaxDivergenceSolverIterations = 100u;
    d->MaxDivergenceSolverError = 10.0f;
    d->EnableViscositySolver = 1;
    d->MinViscositySolverIterations = 0u;
    d->MaxViscositySolverIterations = 100u;
    d->MaxViscositySolverError = 0.1f;
    d->Viscosity = 10.0f;
    d->BoundaryViscosity = 10.0f;
    d->TangentialDistanceFactor = 0.5f;
    d->EnableSurfaceTensionSolver = 1;
    d->SurfaceTensionSmoothPassCount = 1u;
    d->SurfaceTension = 1.0f;
    d->TemporalSmoothing = 0;
    d->CSDFix = -1;
    d->CSD = 10000;
    d->ParticleRadius = 0.025f;
    d->Gravity[0] = 0.0f;
    d->Gravity[1] = -9.81f;
    d->Gravity[2] = 0.0f;
    return VFD_OK;
}

int vfd_batch_create(const vfd_desc* descs, uint32_t n, vfd_sim** out) {
    if (!descs || !out) return fail(VFD_ERROR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    vfd_sim* s = new vfd_sim();
    const int rc = s->impl.init(descs, n);
    if (rc != VFD_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return VFD_OK;
}
int vfd_create(const vfd_desc* desc, vfd_sim** out) { return vfd_batch_create(desc, 1, out); }
int vfd_destroy(vfd_sim* sim) {
    if (sim) {
        sim->impl.join_worker();
        if (sim->impl.stream) (void)hipStreamSynchronize(sim->impl.stream);
        delete sim;
    }
    return VFD_OK;
}
int vfd_set_stream(vfd_sim* sim, void* stream) {
    DRIVE(sim);
    sim->impl.drop_graphs();
    sim->impl.stream = stream ? (hipStream_t)stream : sim->impl.ownStream;
    return VFD_OK;
}
int vfd_scene_count(const vfd_sim* sim, uint32_t* count) {
    CHECK_SIM(sim);
    if (!count) return fail(VFD_ERROR_INVALID_ARGUMENT, "count is NULL");
    *count = (uint32_t)sim->impl.sc.size();
    return VFD_OK;
}
int vfd_batch_set_description(vfd_sim* sim, uint32_t scene, const vfd_desc* desc) {
    DRIVE_SCENE(sim, scene);
    if (!desc) return fail(VFD_ERROR_INVALID_ARGUMENT, "desc is NULL");
    return guard(sim->impl.set_description(scene, *desc));
}
int vfd_set_description(vfd_sim* sim, const vfd_desc* desc) { return vfd_batch_set_description(sim, 0, desc); }
int vfd_get_description(const vfd_sim* sim, vfd_desc* desc) {
    CHECK_SIM(sim);
    if (!desc) return fail(VFD_ERROR_INVALID_ARGUMENT, "desc is NULL");
    std::lock_guard<std::mutex> lk(sim->impl.hostMu);
    *desc = sim->impl.sc[0]->desc;
    return VFD_OK;
}
int vfd_batch_get_info(const vfd_sim* sim, uint32_t scene, vfd_info* info) {
    CHECK_SCENE(sim, scene);
    if (!info) return fail(VFD_ERROR_INVALID_ARGUMENT, "info is NULL");
    std::lock_guard<std::mutex> lk(sim->impl.hostMu);
    const vfd::Scene& S = *sim->impl.sc[scene];
    const vfd::SceneParams& p = S.prm;
    vfd_info& o = *info;
    o.ParticleCount = S.count;
    o.RigidBodyCount = (uint32_t)S.bodies.size();
    o.SupportRadius = p.R;
    o.SupportRadius2 = p.R2;
    o.ParticleRadius = p.r;
    o.ParticleDiameter = p.diam;
    o.TimeStepSize = S.st.dt;
    o.TimeStepSize2 = S.st.dt2;
    o.TimeStepSizeInverse = S.st.dtInv;
    o.TimeStepSize2Inverse = S.st.dt2Inv;
    o.Volume = p.V;
    o.Density0 = p.rho0;
    o.ParticleMass = p.m;
    o.ParticleMassInverse = p.mInv;
    o.Viscosity = S.desc.Viscosity;
    o.BoundaryViscosity = S.desc.BoundaryViscosity;
    o.DynamicViscosity = p.mu;
    o.DynamicBoundaryViscosity = p.muB;
    o.TangentialDistanceFactor = S.desc.TangentialDistanceFactor;
    o.TangentialDistance = p.td;
    o.SurfaceTension = p.sigma;
    o.SurfaceTensionSampleCount = S.st.S;
    o.ClassifierSlope = p.slope;
    o.ClassifierConstant = p.constant;
    o.TemporalSmoothing = (int32_t)p.temporal;
    o.SmoothingFactor = p.smoothing;
    o.Factor = 0.8f;
    o.NeighborParticleRadius = p.NR;
    o.MonteCarloFactor = p.mcFactor;
    o.Gravity[0] = p.gx;
    o.Gravity[1] = p.gy;
    o.Gravity[2] = p.gz;
    return VFD_OK;
}
int vfd_get_info(const vfd_sim* sim, vfd_info* info) { return vfd_batch_get_info(sim, 0, info); }

int vfd_batch_set_fluid(vfd_sim* sim, uint32_t scene, const float* x, const float* v, uint32_t n) {
    DRIVE_SCENE(sim, scene);
    return guard(sim->impl.set_fluid(scene, x, v, n));
}
int vfd_set_fluid(vfd_sim* sim, const float* x, const float* v, uint32_t n) {
    return vfd_batch_set_fluid(sim, 0, x, v, n);
}
int vfd_batch_add_boundary(vfd_sim* sim, uint32_t scene, const vfd_boundary_desc* body) {
    DRIVE_SCENE(sim, scene);
    if (!body) return fail(VFD_ERROR_INVALID_ARGUMENT, "body is NULL");
    return guard(sim->impl.add_boundary(scene, *body));
}
int vfd_add_boundary(vfd_sim* sim, const vfd_boundary_desc* body) { return vfd_batch_add_boundary(sim, 0, body); }
int vfd_batch_add_boundary_mesh(vfd_sim* sim, uint32_t scene, const vfd_mesh_boundary_desc* body,
                                const vfd_mesh* mesh) {
    DRIVE_SCENE(sim, scene);
    if (!body || !mesh) return fail(VFD_ERROR_INVALID_ARGUMENT, "body or mesh is NULL");
    return guard(sim->impl.add_boundary_mesh(scene, *body, *mesh));
}
int vfd_add_boundary_mesh(vfd_sim* sim, const vfd_mesh_boundary_desc* body, const vfd_mesh* mesh) {
    return vfd_batch_add_boundary_mesh(sim, 0, body, mesh);
}
int vfd_clear_boundaries(vfd_sim* sim, uint32_t scene) {
    DRIVE_SCENE(sim, scene);
    return sim->impl.clear_boundaries(scene);
}
int vfd_get_boundary_map(vfd_sim* sim, uint32_t scene, uint32_t body, uint32_t field, float* nodes,
                         uint32_t* node_count) {
    DRIVE_SCENE(sim, scene);
    const vfd::Scene& S = *sim->impl.sc[scene];
    if (body >= S.bodies.size() || field > 1) return fail(VFD_ERROR_INVALID_ARGUMENT, "body or field out of range");
    const vfd::Body& b = *S.bodies[body];
    if (node_count) *node_count = b.map.nodeCount;
    if (nodes) {
        hipError_t e = hipMemcpy(nodes, field == 0 ? b.n0.p : b.n1.p, b.map.nodeCount * sizeof(float),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(VFD_ERROR_HIP, hipGetErrorString(e));
    }
    return VFD_OK;
}

int vfd_step(vfd_sim* sim) {
    DRIVE(sim);
    return guard(sim->impl.step(VFD_PHASE_POSITION));
}
int vfd_step_no_alloc(vfd_sim* sim) {
    DRIVE(sim);
    Simulation& S = sim->impl;
    if (S.capture)
        return fail(VFD_ERROR_STATE, "vfd_step_no_alloc: frame capture allocates frame memory; switch it off with "
                                     "vfd_set_frame_capture(sim, 0)");
    if (S.taping) return fail(VFD_ERROR_STATE, "vfd_step_no_alloc: adjoint recording allocates its records");
    if (S.profiling) return fail(VFD_ERROR_STATE, "vfd_step_no_alloc: profiling creates timing events");
    if (!S.staticScratch) return fail(VFD_ERROR_STATE, "vfd_step_no_alloc: no static scratch on this device");
    if (!S.mc_tables_fit_static())
        return fail(VFD_ERROR_NOT_SUPPORTED, "vfd_step_no_alloc: the surface-tension curvature tables of these "
                                             "descriptions can exceed the static scratch (" +
                                                 std::to_string(vfd::kStaticMcFloats) + " floats, " +
                                                 std::to_string(vfd::kStaticMcScenes) + " scenes)");
    vfd::NoAllocScope scope;
    return guard(S.step(VFD_PHASE_POSITION));
}
int vfd_advance(vfd_sim* sim, uint32_t steps) {
    DRIVE(sim);
    for (uint32_t i = 0; i < steps; ++i) {
        const int rc = sim->impl.step(VFD_PHASE_POSITION);
        if (rc != VFD_OK) return rc;
    }
    return VFD_OK;
}
int vfd_step_until(vfd_sim* sim, int32_t phase) {
    DRIVE(sim);
    if (phase < 0 || phase >= VFD_PHASE_COUNT) return fail(VFD_ERROR_INVALID_ARGUMENT, "phase out of range");
    return guard(sim->impl.step(phase));
}
int vfd_simulate(vfd_sim* sim) {
    DRIVE(sim);
    Simulation& S = sim->impl;
    const int rc = S.committed ? S.upload_baked() : S.commit();
    if (rc != VFD_OK) return rc;
    return S.simulate_frames();
}
int vfd_simulate_async(vfd_sim* sim) {
    CHECK_SIM(sim);
    Simulation& S = sim->impl;
    bool idle = false;
    if (!S.busy.compare_exchange_strong(idle, true))
        return fail(VFD_ERROR_STATE, "the simulation is running on another thread");
    S.join_worker();  // a finished worker nobody waited for
    // baked particles restored and earlier frames released here, before any poll can read them
    const int rc = S.committed ? S.upload_baked() : S.commit();
    if (rc != VFD_OK) {
        S.busy = false;
        return rc;
    }
    S.workerStatus = VFD_OK;
    S.workerError.clear();
    S.simState = VFD_STATE_SIMULATING;
    std::lock_guard<std::mutex> lk(S.workerMu);
    try {
        S.worker = std::thread([&S] {
            int r = hipSetDevice(S.device) == hipSuccess ? VFD_OK : fail(VFD_ERROR_HIP, "hipSetDevice in the worker");
            if (r == VFD_OK) r = S.simulate_frames();
            if (r != VFD_OK) {
                S.workerError = vfd::g_lastError;  // this thread's message, handed to vfd_wait
                S.simState = VFD_STATE_NONE;
            }
            S.workerStatus = r;
            S.busy = false;
        });
    } catch (...) {
        S.simState = VFD_STATE_NONE;
        S.busy = false;
        return fail(VFD_ERROR_STATE, "could not start the simulation thread");
    }
    return VFD_OK;
}
int vfd_wait(vfd_sim* sim) {
    CHECK_SIM(sim);
    const int rc = sim->impl.join_worker();
    return rc == VFD_OK ? VFD_OK : fail(rc, sim->impl.workerError);
}
int vfd_get_simulation_state(const vfd_sim* sim, int32_t* state) {
    CHECK_SIM(sim);
    if (!state) return fail(VFD_ERROR_INVALID_ARGUMENT, "state is NULL");
    *state = sim->impl.simState.load();
    return VFD_OK;
}
int vfd_set_frame_capture(vfd_sim* sim, int32_t enabled) {
    DRIVE(sim);
    sim->impl.capture = enabled != 0;
    return VFD_OK;
}

int vfd_batch_get_particles(vfd_sim* sim, uint32_t scene, float* x, float* v, float* a) {
    DRIVE_SCENE(sim, scene);
    if (x) {
        const int rc = sim->impl.get_field(scene, VFD_FIELD_POSITION, 0, x);
        if (rc != VFD_OK) return rc;
    }
    if (v) {
        const int rc = sim->impl.get_field(scene, VFD_FIELD_VELOCITY, 0, v);
        if (rc != VFD_OK) return rc;
    }
    if (a) {
        const int rc = sim->impl.get_field(scene, VFD_FIELD_ACCELERATION, 0, a);
        if (rc != VFD_OK) return rc;
    }
    return VFD_OK;
}
int vfd_get_particles(vfd_sim* sim, float* x, float* v, float* a) { return vfd_batch_get_particles(sim, 0, x, v, a); }
int vfd_get_field(vfd_sim* sim, uint32_t scene, int32_t field, uint32_t body, float* out) {
    DRIVE_SCENE(sim, scene);
    if (!out) return fail(VFD_ERROR_INVALID_ARGUMENT, "out is NULL");
    return guard(sim->impl.get_field(scene, field, body, out));
}
int vfd_set_field(vfd_sim* sim, uint32_t scene, int32_t field, const float* values) {
    DRIVE_SCENE(sim, scene);
    if (!values) return fail(VFD_ERROR_INVALID_ARGUMENT, "values is NULL");
    return guard(sim->impl.set_field(scene, field, values));
}
int vfd_get_particle_count(const vfd_sim* sim, uint32_t scene, uint32_t* count) {
    CHECK_SCENE(sim, scene);
    if (!count) return fail(VFD_ERROR_INVALID_ARGUMENT, "count is NULL");
    std::lock_guard<std::mutex> lk(sim->impl.hostMu);
    *count = sim->impl.sc[scene]->count;
    return VFD_OK;
}
int vfd_get_neighbors(vfd_sim* sim, uint32_t scene, uint32_t* counts, uint32_t* offsets, uint32_t* ids,
                      uint64_t* total) {
    DRIVE_SCENE(sim, scene);
    return guard(sim->impl.get_neighbors(scene, counts, offsets, ids, total));
}
int vfd_get_debug_info(vfd_sim* sim, uint32_t scene, vfd_debug_info* info) {
    CHECK_SCENE(sim, scene);
    if (!info) return fail(VFD_ERROR_INVALID_ARGUMENT, "info is NULL");
    std::lock_guard<std::mutex> lk(sim->impl.hostMu);
    *info = sim->impl.sc[scene]->dbg;
    info->TimeStepSize = sim->impl.sc[scene]->st.dt;
    info->MaxVelocityMagnitude = sim->impl.sc[scene]->st.maxVel;
    return VFD_OK;
}
int vfd_get_frame_count(vfd_sim* sim, uint32_t scene, uint32_t* count) {
    CHECK_SCENE(sim, scene);
    if (!count) return fail(VFD_ERROR_INVALID_ARGUMENT, "count is NULL");
    std::lock_guard<std::mutex> lk(sim->impl.hostMu);
    *count = (uint32_t)sim->impl.sc[scene]->frames.size();
    return VFD_OK;
}
int vfd_get_frame(vfd_sim* sim, uint32_t scene, uint32_t frame, float* xva, float* max_velocity, float* time_step) {
    CHECK_SCENE(sim, scene);
    Simulation& S = sim->impl;
    // the frame's asynchronous copy must have landed; wait for it without holding the lock that a running step's
    // bookkeeping takes, then copy out under the lock (frames are released only by driving calls)
    for (;;) {
        hipEvent_t done = nullptr;
        {
            std::lock_guard<std::mutex> lk(S.hostMu);
            const auto& fr = S.sc[scene]->frames;
            if (frame >= fr.size()) return fail(VFD_ERROR_INVALID_ARGUMENT, "frame index out of range");
            const vfd::Frame& f = fr[frame];
            if (!f.done || hipEventQuery(f.done) == hipSuccess) {
                if (xva) std::memcpy(xva, f.data(), f.n * sizeof(float));
                if (max_velocity) *max_velocity = f.maxVel;
                if (time_step) *time_step = f.dt;
                return VFD_OK;
            }
            done = f.done;
        }
        (void)hipGetLastError();  // the query's "not ready" is no error of this thread
        VFD_HIP(hipEventSynchronize(done));
    }
}
int vfd_set_profiling(vfd_sim* sim, int32_t enabled) {
    DRIVE(sim);
    sim->impl.profiling = enabled != 0;
    sim->impl.kstats.clear();
    sim->impl.pending.clear();
    sim->impl.evUsed = 0;
    return VFD_OK;
}
int vfd_get_profile(vfd_sim* sim, vfd_kernel_stat* out, uint32_t capacity, uint32_t* count) {
    DRIVE(sim);
    const auto& ks = sim->impl.kstats;
    if (count) *count = (uint32_t)ks.size();
    if (out)
        for (uint32_t i = 0; i < capacity && i < ks.size(); ++i) {
            std::memset(&out[i], 0, sizeof(vfd_kernel_stat));
            std::strncpy(out[i].name, ks[i].name.c_str(), sizeof(out[i].name) - 1);
            out[i].launches = ks[i].launches;
            out[i].total_ms = ks[i].ms;
        }
    return VFD_OK;
}
int vfd_tape(vfd_sim* sim, int32_t on) {
    DRIVE(sim);
    auto& S = sim->impl;
    S.taping = on != 0;  // (solver chunks run without their graphs while recording)
    if (on) {
        S.tapeCount = 0;
        S.tapeError.clear();
    }
    return VFD_OK;
}
int vfd_tape_steps(vfd_sim* sim, uint32_t* steps) {
    DRIVE(sim);
    if (!steps) return fail(VFD_ERROR_INVALID_ARGUMENT, "steps must not be NULL");
    *steps = (uint32_t)sim->impl.tapeCount;
    return VFD_OK;
}
int vfd_adjoint(vfd_sim* sim, const float* lx_out, const float* lv_out, float* lx_in, float* lv_in, double* lmu) {
    DRIVE(sim);
    if (!lx_out || !lx_in || !lv_in) return fail(VFD_ERROR_INVALID_ARGUMENT, "cotangent buffers must not be NULL");
    return sim->impl.adjoint(0, lx_out, lv_out, lx_in, lv_in, lmu);
}
int vfd_batch_adjoint(vfd_sim* sim, uint32_t scene, const float* lx_out, const float* lv_out, float* lx_in,
                      float* lv_in, double* lmu) {
    DRIVE_SCENE(sim, scene);
    if (!lx_out || !lx_in || !lv_in) return fail(VFD_ERROR_INVALID_ARGUMENT, "cotangent buffers must not be NULL");
    return sim->impl.adjoint(scene, lx_out, lv_out, lx_in, lv_in, lmu);
}
int vfd_get_tuning(vfd_sim* sim, vfd_tuning* out) {
    DRIVE(sim);
    if (!out) return fail(VFD_ERROR_INVALID_ARGUMENT, "out must not be NULL");
    out->SolverCheckInterval = sim->impl.pollChunk;
    out->HaloLimit = sim->impl.haloLimit;
    out->UseGraphs = sim->impl.useGraphs ? 1u : 0u;
    out->PredictIterations = sim->impl.predictIterations ? 1u : 0u;
    out->AdjointGroupLoops = sim->impl.adjointGroups;
    return VFD_OK;
}
int vfd_set_tuning(vfd_sim* sim, const vfd_tuning* t) {
    DRIVE(sim);
    if (!t) return fail(VFD_ERROR_INVALID_ARGUMENT, "tuning must not be NULL");
    if (t->SolverCheckInterval < 1 || t->SolverCheckInterval > 1024)
        return fail(VFD_ERROR_INVALID_ARGUMENT, "SolverCheckInterval must be in [1, 1024]");
    if (t->HaloLimit > vfd::kHaloMax)
        return fail(VFD_ERROR_INVALID_ARGUMENT, "HaloLimit must be at most " + std::to_string(vfd::kHaloMax));
    if (t->AdjointGroupLoops > 2) return fail(VFD_ERROR_INVALID_ARGUMENT, "AdjointGroupLoops must be 0, 1 or 2");
    if (t->SolverCheckInterval != sim->impl.pollChunk) sim->impl.drop_graphs();
    sim->impl.pollChunk = t->SolverCheckInterval;
    sim->impl.haloLimit = t->HaloLimit;
    sim->impl.useGraphs = t->UseGraphs != 0;
    sim->impl.predictIterations = t->PredictIterations != 0;
    sim->impl.adjointGroups = t->AdjointGroupLoops;
    return VFD_OK;
}
int vfd_get_kernel_table(vfd_sim* sim, float* w, float* gradient_w, float* scalars) {
    DRIVE(sim);
    const vfd::KernelTable& k = sim->impl.kt;
    if (w) std::memcpy(w, k.W.data(), k.W.size() * sizeof(float));
    if (gradient_w) std::memcpy(gradient_w, k.G.data(), k.G.size() * sizeof(float));
    if (scalars) {
        scalars[0] = k.R;
        scalars[1] = k.R2;
        scalars[2] = k.invStep;
        scalars[3] = k.W0;
        scalars[4] = k.K;
        scalars[5] = k.L;
    }
    return VFD_OK;
}

}  // extern "C"
