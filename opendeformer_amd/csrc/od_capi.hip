// C-ABI of libodhip.so (declared in include/od.h): handles, device memory,
// PCG driver (FP64 and FP32-inner / FP64-refinement) and the integrators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/od.h"
#include "od_internal.h"
#include "od_precond.h"

using namespace od;

namespace {
thread_local std::string g_err;

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string &s) : std::runtime_error(s), code(c) {}
};

#define API_BEGIN try {
#define API_END                                                        \
    }                                                                  \
    catch (const od::HipError &e) { g_err = e.what(); return OD_ERR_HIP; } \
    catch (const ApiError &e) { g_err = e.what(); return e.code; }     \
    catch (const std::exception &e) { g_err = e.what(); return OD_ERR_ARG; } \
    return OD_OK;

void require(bool c, const char *msg, int code = OD_ERR_ARG) {
    if (!c) throw ApiError(code, msg);
}

__global__ void k_done_to_mask(int n, const InstScalars *sc, const uint8_t *extra, uint8_t *mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mask[i] = (sc[i].done || (extra && extra[i])) ? 1 : 0;
}

// ---- step graph helpers (a captured graph holds no memset node: these kernels clear and copy)
__global__ void k_zero_words(uint32_t *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void k_copy_f64(double *dst, const double *src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
// iterations of one refinement sweep, per instance
__global__ void k_record_sweep(int n, const InstScalars *sc, int *rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rec[i] = sc[i].it;
}
// The step graph commits an instance when its solve is done as the eager step would leave it:
// FP64, converged or at maxIter within the graph's iterations; mixed, the outer residual
// converged with no refinement sweep stopped by the graph's iteration count.
struct SweepPlan {
    int n;
    int k[16];
};
__global__ void k_commit_flags(int n, const InstScalars *sc, const int *rec, SweepPlan plan, const uint8_t *active,
                               uint8_t *commit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool ok = sc[i].done && (!active || active[i]);
    if (rec)
        for (int q = 0; q < plan.n; q++) ok = ok && rec[q * n + i] < plan.k[q];
    commit[i] = ok ? 1 : 0;
}
// active instances the graph did not commit
__global__ void k_uncommitted(int n, const uint8_t *active, const uint8_t *commit, uint8_t *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ((!active || active[i]) && !commit[i]) ? 1 : 0;
}
}  // namespace

// ---- fixed scratch of od_pcg_solve_noalloc ---------------------------------------------
// A file-scope device array (one copy per GPU, held until the process ends) that carries all
// the temporaries of the allocation-free PCG solve.  Host code never passes the array's name
// to a kernel (on the host it is only a stand-in symbol): the device address of each GPU's
// copy is looked up once per device with hipGetSymbolAddress when a context is created on it,
// and that address is what the solve hands to its kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::atomic<unsigned char *> g_scratchDev[kMaxDevices];

// Called with `device` current.
static void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices || g_scratchDev[device].load()) return;
    void *p = nullptr;
    OD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchDev[device].store((unsigned char *)p);
}

struct od_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // meshes bound to this context; od_ctx_destroy with meshes still bound defers the release
    // of the stream until the last of them is destroyed or rebound
    std::atomic<int> meshes{0};
    std::atomic<bool> closed{false};
    int *pinned = nullptr;   // host-pinned mailbox for convergence polling
    int pinnedCount = 0;
    // optional HIP-event timing of the PCG SpMV launches (measurement mode)
    bool prof = false;
    std::vector<hipEvent_t> ev;
    size_t evUsed = 0;
    double profMs = 0.0;
    long long profCount = 0;
};

// Accumulates the elapsed time of completed event pairs (caller has synchronised the stream).
static void prof_flush(od_ctx *c) {
    for (size_t k = 0; k + 1 < c->evUsed; k += 2) {
        float ms = 0;
        OD_HIP(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
        c->profMs += ms;
        c->profCount++;
    }
    c->evUsed = 0;
}
// Next (start, stop) event pair for a timed launch, or nulls outside measurement mode.
static void prof_pair(od_ctx *c, hipEvent_t *e0, hipEvent_t *e1) {
    *e0 = *e1 = nullptr;
    if (!c->prof) return;
    if (c->evUsed + 2 > c->ev.size()) {
        OD_HIP(hipEventSynchronize(c->ev[c->evUsed - 1]));
        prof_flush(c);
    }
    *e0 = c->ev[c->evUsed++];
    *e1 = c->ev[c->evUsed++];
}

struct od_mesh {
    od_ctx *ctx = nullptr;
    // references: the handle itself and every solver, matrix handle and integrator built on the
    // mesh; the last release frees it, so destroying them in any order is safe
    std::atomic<int> refs{1};
    Topology topo;
    DevMesh dm{};
    std::vector<void *> allocs;
    size_t bytes = 0;
    bool hasMaterial = false;
    bool spmvTuned = false;                      // SpMV variant chosen (autotuner or caller)
    // generation of the launch forms (SpMV shapes, sweep form per level): bumped by every
    // change, part of the key of the captured PCG graphs, which bake the forms in
    uint64_t formGen = 0;
    double tuneMs[2][3][3] = {};   // [FP64, FP32][16, 8, 4 lanes per row][R = 1, 2, 4]: cold launch time
    double tuneSvMs[3] = {};       // FP32 slot-value form, [1, 2, 4 slots per thread]: cold launch time
    double tuneGrpMs[2][2] = {{-1, -1}, {-1, -1}};   // grouped form, [FP64, FP32] x [4, 8 slots per thread]: cold launch time (-1: not available)
    double *blkC = nullptr, *nodeC = nullptr;   // contribution arrays (lazily allocated)
    // Partially constrained nodes: the caller's DOF vectors hold topo.ndofFree entries, the
    // kernels' topo.ndof (the last block row padded).  Entry points that take device vectors
    // run on padded copies kept here (zeroed once; the kernels leave the padding at zero), and
    // upload/download map the reference's value layout (padDiag: the padding diagonal values).
    void *padBuf[2][3] = {};                     // [double, float][slot]
    std::vector<int64_t> padDiag;
    template <typename T>
    T *padded(int slot) {
        void *&b = padBuf[sizeof(T) == 4][slot];
        if (!b) {
            b = alloc(sizeof(T) * (size_t)topo.ndof);
            OD_HIP(hipMemsetAsync(b, 0, sizeof(T) * (size_t)topo.ndof, ctx->stream));
        }
        return (T *)b;
    }
    std::vector<uint8_t> levelLanesFwd, levelLanesBwd;       // DevMesh points at these (host memory)
    std::vector<uint8_t> levelLanesFwd64, levelLanesBwd64;   // the same for the FP64 sweeps
    std::vector<int> levelChunkPtr;                      // per level, its range in dm.levelChunks
    template <typename T>
    T *upload(const std::vector<T> &v) {
        if (v.empty()) return nullptr;
        void *p = nullptr;
        OD_HIP(hipMalloc(&p, v.size() * sizeof(T)));
        allocs.push_back(p);
        bytes += v.size() * sizeof(T);
        OD_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
        return (T *)p;
    }
    void *alloc(size_t b) {
        void *p = nullptr;
        OD_HIP(hipMalloc(&p, std::max<size_t>(b, 16)));
        allocs.push_back(p);
        bytes += b;
        return p;
    }
    void ensureContrib() {
        if (!blkC) blkC = (double *)alloc(sizeof(double) * 9 * std::max<int64_t>(1, topo.nPairContrib));
        if (!nodeC) nodeC = (double *)alloc(sizeof(double) * 3 * std::max<int64_t>(1, topo.nNodeContrib));
    }
    // plane collision damping: per-block counters (zeroed once, left at zero by every call)
    // and per-pair buffers grown to the largest call so far
    DampScratch damp;
    int dampCap = 0;
    std::vector<void *> dampPairAllocs;
    void ensureDamp(int nPairs) {
        if (!damp.cnt) {
            const size_t b = sizeof(int) * (size_t)std::max<int64_t>(1, topo.nBlocks);
            damp.cnt = (int *)alloc(b);
            damp.fill = (int *)alloc(b);
            damp.start = (int *)alloc(b);
            damp.nTouched = (int *)alloc(sizeof(int));
            OD_HIP(hipMemsetAsync(damp.cnt, 0, b, ctx->stream));
            OD_HIP(hipMemsetAsync(damp.fill, 0, b, ctx->stream));
        }
        if (nPairs <= dampCap) return;
        OD_HIP(hipStreamSynchronize(ctx->stream));
        for (void *p : dampPairAllocs) (void)hipFree(p);
        dampPairAllocs.clear();
        const int cap = std::max(nPairs, 2 * dampCap);
        auto get = [&](size_t b) {
            void *p = nullptr;
            OD_HIP(hipMalloc(&p, b));
            dampPairAllocs.push_back(p);
            return p;
        };
        damp.planes = (double *)get(sizeof(double) * 4 * (size_t)cap);
        damp.tets = (int *)get(sizeof(int) * (size_t)cap);
        damp.pv = (double *)get(sizeof(double) * 90 * (size_t)cap);
        damp.dst = (int *)get(sizeof(int) * 10 * (size_t)cap);
        damp.list = (int *)get(sizeof(int) * 10 * (size_t)cap);
        damp.touched = (int *)get(sizeof(int) * 10 * (size_t)cap);
        dampCap = cap;
    }
    ~od_mesh() {
        for (void *p : dampPairAllocs) (void)hipFree(p);
        for (void *p : allocs) (void)hipFree(p);
    }
};

// A caller's device DOF vector as the kernels see it: the vector itself, or (partially
// constrained nodes) a padded copy, filled from it when `in` and copied back by done() when
// `out`.  Calls on one mesh are ordered on its context stream, so the mesh's copies serve them
// in turn.
template <typename T>
struct DofVec {
    od_mesh *mesh;
    T *user;
    T *dev;
    bool out;
    DofVec(od_mesh *m, int slot, const T *u, bool in, bool outp) : mesh(m), user(const_cast<T *>(u)), dev(user), out(outp) {
        if (!m->topo.partial || !u) return;
        dev = m->padded<T>(slot);
        if (in)
            OD_HIP(hipMemcpyAsync(dev, u, sizeof(T) * (size_t)m->topo.ndofFree, hipMemcpyDeviceToDevice, m->ctx->stream));
    }
    void done() {
        if (dev != user && out)
            OD_HIP(hipMemcpyAsync(user, dev, sizeof(T) * (size_t)mesh->topo.ndofFree, hipMemcpyDeviceToDevice,
                                  mesh->ctx->stream));
    }
};

// reference value layout <-> the kernels' (identity unless partially constrained)
static void values_to_kernel(const od_mesh *mesh, const double *ref, std::vector<double> &out) {
    const Topology &t = mesh->topo;
    out.assign((size_t)t.nValues, 0.0);
    for (int64_t v = 0; v < t.nValues; v++)
        if (t.refPos[v] >= 0) out[v] = ref[t.refPos[v]];
    for (int64_t v : mesh->padDiag) out[v] = 1.0;
}
static void values_from_kernel(const od_mesh *mesh, const double *in, double *ref) {
    const Topology &t = mesh->topo;
    for (int64_t v = 0; v < t.nValues; v++)
        if (t.refPos[v] >= 0) ref[t.refPos[v]] = in[v];
}

static void mesh_retain(od_mesh *mesh);
static void mesh_release(od_mesh *mesh);

struct od_matrix {
    od_mesh *mesh = nullptr;
    bool retained = false;   // holds a mesh reference (a handle; not a matrix inside an integrator)
    double *vals = nullptr;
    float *vals32 = nullptr;
    ~od_matrix() {
        if (vals) (void)hipFree(vals);
        if (vals32) (void)hipFree(vals32);
        if (retained) mesh_release(mesh);
    }
};

// The reference's scalar preconditioners (threshold IC, SAINV LDL^T): factored on the host at
// every solve from the downloaded values, swept on the device.
struct ScalarPrecond {
    ScalarFactor factor;
    ScalarSweepsHost host;
    ScalarSweepsDev dev;
    std::vector<double> values;              // host copy of A's values
    void *buf = nullptr;                     // device arrays of the sweeps, one allocation
    size_t cap = 0;
    ~ScalarPrecond() { if (buf) (void)hipFree(buf); }
};

struct od_solver {
    od_mesh *mesh = nullptr;
    bool retained = false;      // holds a mesh reference (not the no-allocation solve's view)
    int precond = OD_PRE_JACOBI;
    double param = 1e-3;        // IC(0) pivot epsilon; threshold IC epsilon; SAINV epsilon
    double param2 = 1.0;        // threshold IC relaxedScale; LDL^T entry epsilon
    ScalarPrecond *scalar = nullptr;
    // the preconditioner holds a factor of a matrix (od_precond_preprocess / od_precond_apply;
    // a PCG solve rebuilds it from its own matrix)
    bool preprocessed = false;
    int precision = OD_FP64;
    double innerTol = 1e-4;     // mixed: inner FP32 relative tolerance (floor; see kInnerOuterScale2)
    int maxOuter = 12;
    std::vector<void *> allocs;
    double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *invDiag = nullptr, *U = nullptr;
    double *U64 = nullptr, *Ul64 = nullptr;   // IC(0): the FP64 sweeps' part-major copies of U
    float *r32 = nullptr, *ri32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
    float *invDiag32 = nullptr, *U32 = nullptr, *Ul32 = nullptr;
    float *sv32 = nullptr;      // FP32 slot values of the matrix (slot-value product form)
    double *gv64 = nullptr;     // grouped values of the matrix (grouped product form), FP64
    float *gv32 = nullptr;      // and FP32
    double *Ax = nullptr;
    double *partials = nullptr;
    int *doneIn = nullptr, *doneOut = nullptr;   // per-instance done flags (device)
    InstScalars *scIn = nullptr, *scOut = nullptr;
    uint8_t *mask = nullptr, *zeroMask = nullptr;
    std::vector<int> iters;
    std::vector<double> relres;
    int outer = 0;
    int lastMaxIters = 0;       // iterations of the previous solve (sizes the first launch batch)
    std::vector<int> sweepIters; // mixed form: iterations of each inner sweep of the previous solve
    size_t bytes = 0;
    // Zero-filled in the mesh's stream order (the first product of a solve forms p = z + 0 *
    // p_old, which a stale NaN or infinity in a fresh direction buffer would poison).  A
    // synchronous hipMemset runs on the null stream, which the non-blocking context streams do
    // not wait for: under concurrent streams it could land after the kernels that fill the buffer.
    void *alloc(size_t b) {
        void *ptr = nullptr;
        OD_HIP(hipMalloc(&ptr, std::max<size_t>(b, 16)));
        allocs.push_back(ptr);
        OD_HIP(hipMemsetAsync(ptr, 0, std::max<size_t>(b, 16), mesh->ctx->stream));
        bytes += b;
        return ptr;
    }
    // graph mode: PCG iterations replayed from a captured block of kGraphIters iterations
    bool useGraph = false;
    struct IterGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<const void *> key;   // the pointers baked into the captured launches
    } graph[2];                          // [FP64, FP32]
    long long captures = 0;              // graph captures so far (each capture instantiates a graph)
    int *sweepRec = nullptr;             // step graph, mixed form: iterations per (sweep, instance)
    ~od_solver() {
        for (auto &g : graph)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (void *ptr : allocs) (void)hipFree(ptr);
        delete scalar;
        if (retained) mesh_release(mesh);
    }
};

// Even: the FP32 iterations alternate between two direction buffers by iteration parity,
// and a replayed block must start on the parity it was captured with.
constexpr int kGraphIters = 4;
static_assert(kGraphIters % 2 == 0, "graph blocks hold an even number of iterations");
// Graph mode captures a block only while it holds at most this many kernel launches.  An
// IC(0) iteration launches about 2 x (dependency levels) sweep kernels; a numbering with
// thousands of levels (a bandwidth-reducing relabelling of a large mesh) would make a block
// of tens of thousands of nodes, whose capture and instantiation cost more than the launch
// overhead the replay saves on launches that each run a level's worth of rows.
constexpr int kGraphMaxLaunches = 4096;

struct od_integrator {
    od_mesh *mesh = nullptr;
    bool retained = false;   // holds a mesh reference
    int kind = 0;  // 0 Newmark, 1 backward Euler
    double beta = 0.25, gamma = 0.5, h = 1.0 / 60, alphaM = 0, alphaK = 0, tol = 1e-6;
    int maxIter = 10000;
    bool initAccel = true;
    int steps = 0;
    od_solver *solver = nullptr;
    od_matrix M, A, Dm;   // Dm: the damping matrix D as a handle (backward Euler, od_integrator_damping_matrix)
    double *d = nullptr, *v = nullptr, *a = nullptr, *aNew = nullptr, *fext = nullptr, *imp = nullptr, *rhs = nullptr;
    uint8_t *active = nullptr, *inactive = nullptr;
    bool hasImpulse = false, hasActive = false;
    // external damping matrix (backward Euler; allocated on the first add) and D v
    double *D = nullptr, *Dv = nullptr;
    bool hasDamp = false;
    std::vector<void *> allocs;
    std::vector<int> lastIters;
    std::vector<double> lastRelres;
    int lastOuter = 0, initIters = 0;
    size_t bytes = 0;
    // Whole-step graphs (graph mode, Newmark): element -> gather -> factor -> a fixed number of
    // PCG iterations (or refinement sweeps) with device-side exits -> commit, one launch per
    // step, cached by iteration plan.
    struct StepGraph {
        std::vector<int> key;
        hipGraphExec_t exec = nullptr;
        uint64_t lastUse = 0;
    };
    std::vector<StepGraph> stepGraphs;
    uint64_t stepClock = 0;
    uint8_t *commit = nullptr, *redo = nullptr;   // per instance: committed by the graph; to complete eagerly
    long long graphSteps = 0, graphFallbacks = 0, graphCaptures = 0;
    bool stepGraph = false;   // graph mode 2: whole-step graphs (mode 1: PCG iteration graphs only)
    // zero-filled in the mesh's stream order (the state starts at rest; see od_solver::alloc)
    void *alloc(size_t b) {
        void *ptr = nullptr;
        OD_HIP(hipMalloc(&ptr, std::max<size_t>(b, 16)));
        allocs.push_back(ptr);
        OD_HIP(hipMemsetAsync(ptr, 0, std::max<size_t>(b, 16), mesh->ctx->stream));
        bytes += b;
        return ptr;
    }
    ~od_integrator() {
        for (auto &g : stepGraphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        M.vals = nullptr; A.vals = nullptr; A.vals32 = nullptr; Dm.vals = nullptr;   // owned by allocs
        for (void *ptr : allocs) (void)hipFree(ptr);
        delete solver;
        if (retained) mesh_release(mesh);
    }
};

// =====================================================================================
// solver internals
// =====================================================================================

// Autotuner of the PCG SpMV: times each chunks-per-workgroup variant (1, 2, 4) for FP64 and
// FP32 values on this mesh's pattern, with the matrix out of every cache as it is inside a
// PCG iteration (a 512 MB streaming read before each launch evicts the 256 MB memory-side
// cache), and keeps the fastest median of three.  Values are zero: the time depends only on
// the pattern.
static void autotune_spmv(od_mesh *mesh) {
    if (mesh->spmvTuned) return;
    DevMesh &m = mesh->dm;
    hipStream_t st = mesh->ctx->stream;
    const Topology &t = mesh->topo;
    const size_t nv = (size_t)t.nValues + kValuePad, n = (size_t)t.ndof;
    const int64_t flushN = (int64_t)128 << 20;    // floats: 512 MB
    std::vector<void *> bufs;
    auto get = [&](size_t bytes) {
        void *p = nullptr;
        OD_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        bufs.push_back(p);
        OD_HIP(hipMemsetAsync(p, 0, std::max<size_t>(bytes, 16), st));
        return p;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        // the PCG products as the solver runs them: FP32 with the fused direction update (beta = 0)
        void *vals = get(8 * nv), *x = get(8 * n), *y = get(8 * n), *z = get(8 * n), *pNew = get(8 * n);
        double *partials = (double *)get(8 * (size_t)std::max(1, m.nChunks));
        InstScalars *sc = (InstScalars *)get(sizeof(InstScalars) * t.nInst);   // zero: no instance done
        float *flush = (float *)get(4 * (size_t)flushN);
        float *sv = (float *)get(sizeof(float) * kSlotValFloats * (size_t)std::max(1, m.nChunks));
        void *gvals = (m.grpOk64 || m.grpOk32) ? get(8 * (size_t)grp_values(m)) : nullptr;
        OD_HIP(hipEventCreate(&e0));
        OD_HIP(hipEventCreate(&e1));
        // the grouped product (FP64 plain, FP32 with the fused direction update)
        for (int prec = 0; prec < 2; prec++)
            for (int kv = 0; kv < 2; kv++) {
                const int K = kGrpSlotVariants[kv];
                mesh->tuneGrpMs[prec][kv] = -1.0;
                if (!(prec == 0 ? m.grpOk64 : m.grpOk32)) continue;
                float ms[3];
                for (int rep = 0; rep < 3; rep++) {
                    launch_stream_read(st, flush, flushN);
                    if (prec == 0)
                        launch_spmv_dot<double>(m, st, K, -1, (const double *)vals, (const double *)x, (double *)y,
                                                partials, sc, e0, e1, nullptr, nullptr, (const double *)gvals);
                    else
                        launch_spmv_dot<float>(m, st, K, -1, (const float *)vals, (const float *)x, (float *)y,
                                               partials, sc, e0, e1, (const float *)z, (float *)pNew,
                                               (const float *)gvals);
                    OD_HIP(hipEventSynchronize(e1));
                    OD_HIP(hipEventElapsedTime(&ms[rep], e0, e1));
                }
                std::sort(ms, ms + 3);
                mesh->tuneGrpMs[prec][kv] = ms[1];
            }
        for (int v = 0; v < 3; v++) {
            float ms[3];
            for (int rep = 0; rep < 3; rep++) {
                launch_stream_read(st, flush, flushN);
                launch_spmv_dot<float>(m, st, kSpmvVariants[v], 0, (const float *)vals, (const float *)x, (float *)y,
                                       partials, sc, e0, e1, (const float *)z, (float *)pNew, sv);
                OD_HIP(hipEventSynchronize(e1));
                OD_HIP(hipEventElapsedTime(&ms[rep], e0, e1));
            }
            std::sort(ms, ms + 3);
            mesh->tuneSvMs[v] = ms[1];
        }
        for (int prec = 0; prec < 2; prec++)
            for (int lv = 0; lv < 3; lv++)
                for (int v = 0; v < 3; v++) {
                    const int R = kSpmvVariants[v], LS = kSpmvLaneVariants[lv];
                    float ms[3];
                    for (int rep = 0; rep < 3; rep++) {
                        launch_stream_read(st, flush, flushN);
                        if (prec == 0)
                            launch_spmv_dot<double>(m, st, R, LS, (const double *)vals, (const double *)x, (double *)y,
                                                    partials, sc, e0, e1);
                        else
                            launch_spmv_dot<float>(m, st, R, LS, (const float *)vals, (const float *)x, (float *)y,
                                                   partials, sc, e0, e1, (const float *)z, (float *)pNew);
                        OD_HIP(hipEventSynchronize(e1));
                        OD_HIP(hipEventElapsedTime(&ms[rep], e0, e1));
                    }
                    std::sort(ms, ms + 3);
                    mesh->tuneMs[prec][lv][v] = ms[1];
                }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void *p : bufs) (void)hipFree(p);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (void *p : bufs) (void)hipFree(p);
    int best[2][2] = {{0, 0}, {0, 0}};   // [precision] -> (lane variant, R variant)
    for (int prec = 0; prec < 2; prec++)
        for (int lv = 0; lv < 3; lv++)
            for (int v = 0; v < 3; v++)
                if (mesh->tuneMs[prec][lv][v] < mesh->tuneMs[prec][best[prec][0]][best[prec][1]]) {
                    best[prec][0] = lv;
                    best[prec][1] = v;
                }
    m.spmvR64 = m.spmvR64d = kSpmvVariants[best[0][1]];
    m.spmvL64 = m.spmvL64d = kSpmvLaneVariants[best[0][0]];
    m.spmvR32 = kSpmvVariants[best[1][1]];
    m.spmvL32 = kSpmvLaneVariants[best[1][0]];
    // the FP32 slot-value form (lanes 0) when one of its shapes is faster still
    double best32 = mesh->tuneMs[1][best[1][0]][best[1][1]];
    for (int v = 0; v < 3; v++)
        if (mesh->tuneSvMs[v] < best32) {
            best32 = mesh->tuneSvMs[v];
            m.spmvR32 = kSpmvVariants[v];
            m.spmvL32 = 0;
        }
    // the grouped form (lanes -1, R = slots per thread) when one of its shapes is faster still
    double best64 = mesh->tuneMs[0][best[0][0]][best[0][1]];
    for (int kv = 0; kv < 2; kv++) {
        if (mesh->tuneGrpMs[0][kv] >= 0 && mesh->tuneGrpMs[0][kv] < best64) {
            best64 = mesh->tuneGrpMs[0][kv];
            m.spmvR64 = kGrpSlotVariants[kv];
            m.spmvL64 = -1;
        }
        if (mesh->tuneGrpMs[1][kv] >= 0 && mesh->tuneGrpMs[1][kv] < best32) {
            best32 = mesh->tuneGrpMs[1][kv];
            m.spmvR32 = kGrpSlotVariants[kv];
            m.spmvL32 = -1;
        }
    }
    // lanes per row of the IC(0) sweeps, per level (meshes with few levels)
    if (t.nLevels <= 64) {
        tune_ic0_lanes<float>(m, st, t.levelPtr, mesh->levelLanesFwd.data(), mesh->levelLanesBwd.data());
        tune_ic0_lanes<double>(m, st, t.levelPtr, mesh->levelLanesFwd64.data(), mesh->levelLanesBwd64.data());
    }
    mesh->spmvTuned = true;
    mesh->formGen++;
}

static bool scalar_precond(int precond) { return precond == OD_PRE_ICT || precond == OD_PRE_LDLT; }

// The FP64 sweeps' copies of the IC(0) factor (once per solver).
static void alloc_fp64_sweeps(od_solver *s) {
    if (s->U64) return;
    const DevMesh &m = s->mesh->dm;
    s->U64 = (double *)s->alloc(sizeof(double) * (size_t)ic0_u32_floats(m));
    s->Ul64 = (double *)s->alloc(sizeof(double) * (size_t)std::max<int64_t>(1, ic0_ul32_floats(m)));
}

static od_solver *make_solver(od_mesh *mesh, int precond, double param, int precision) {
    require(precond == OD_PRE_NONE || precond == OD_PRE_JACOBI || precond == OD_PRE_IC0 || scalar_precond(precond),
            "unknown preconditioner", OD_ERR_UNSUPPORTED);
    require(precision == OD_FP64 || precision == OD_MIXED, "unknown precision", OD_ERR_UNSUPPORTED);
    require(!(scalar_precond(precond) && precision == OD_MIXED),
            "the threshold IC and LDL^T preconditioners run in FP64 only", OD_ERR_UNSUPPORTED);
    autotune_spmv(mesh);
    od_solver *s = new od_solver();
    try {
        s->mesh = mesh;
        mesh_retain(mesh);
        s->retained = true;
        s->precond = precond;
        s->param = scalar_precond(precond) ? param : (param > 0 ? param : 1e-3);
        s->param2 = precond == OD_PRE_LDLT ? 0.0 : 1.0;
        s->precision = precision;
        if (scalar_precond(precond)) s->scalar = new ScalarPrecond();
        s->sweepIters.assign(s->maxOuter, 0);
        const Topology &t = mesh->topo;
        size_t n = (size_t)t.ndof, nv = (size_t)t.nValues;
        s->r = (double *)s->alloc(8 * n); s->z = (double *)s->alloc(8 * n);
        s->p = (double *)s->alloc(8 * n); s->q = (double *)s->alloc(8 * n);
        s->Ax = (double *)s->alloc(8 * n);
        if (precond == OD_PRE_JACOBI) s->invDiag = (double *)s->alloc(8 * n);
        if (precond == OD_PRE_IC0) {
            s->U = (double *)s->alloc(8 * (nv + kValuePad));
            // the FP64 sweep copies serve only the FP64 sweeps: mixed solvers allocate them on
            // the first od_precond_apply
            if (precision == OD_FP64) alloc_fp64_sweeps(s);
        }
        if (precision == OD_MIXED) {
            s->r32 = (float *)s->alloc(4 * n); s->ri32 = (float *)s->alloc(4 * n); s->z32 = (float *)s->alloc(4 * n);
            s->p32 = (float *)s->alloc(4 * 2 * n); s->q32 = (float *)s->alloc(4 * n);   // two direction buffers
            s->x32 = (float *)s->alloc(4 * n);
            if (precond == OD_PRE_JACOBI) s->invDiag32 = (float *)s->alloc(4 * n);
            if (precond == OD_PRE_IC0) {
                s->U32 = (float *)s->alloc(4 * (size_t)ic0_u32_floats(mesh->dm));
                s->Ul32 = (float *)s->alloc(4 * (size_t)std::max<int64_t>(1, ic0_ul32_floats(mesh->dm)));
            }
            // the slot values of the FP32 product, when the mesh runs that form (allocated
            // here so the solver's device bytes include them; a later change of form
            // allocates them at the next solve)
            if (mesh->dm.spmvL32 == 0)
                s->sv32 = (float *)s->alloc(sizeof(float) * kSlotValFloats * (size_t)std::max(1, mesh->dm.nChunks));
            if (mesh->dm.spmvL32 == -1) s->gv32 = (float *)s->alloc(sizeof(float) * (size_t)grp_values(mesh->dm));
        }
        if (precision == OD_FP64 && mesh->dm.spmvL64 == -1)
            s->gv64 = (double *)s->alloc(sizeof(double) * (size_t)grp_values(mesh->dm));
        s->partials = (double *)s->alloc(sizeof(double) * 3 * (size_t)t.chunkRow0.size());
        s->doneIn = (int *)s->alloc(sizeof(int) * t.nInst);
        s->doneOut = (int *)s->alloc(sizeof(int) * t.nInst);
        s->scIn = (InstScalars *)s->alloc(sizeof(InstScalars) * t.nInst);
        s->scOut = (InstScalars *)s->alloc(sizeof(InstScalars) * t.nInst);
        s->mask = (uint8_t *)s->alloc(t.nInst);
        s->zeroMask = (uint8_t *)s->alloc(t.nInst);
        s->iters.assign(t.nInst, 0);
        s->relres.assign(t.nInst, 0.0);
    } catch (...) {
        delete s;
        throw;
    }
    return s;
}

// Number of instances whose done flag is still 0 (synchronises the stream).
static int count_not_done(od_ctx *ctx, const int *doneDev, int nInst) {
    if (ctx->pinnedCount < nInst) {
        if (ctx->pinned) OD_HIP(hipHostFree(ctx->pinned));
        OD_HIP(hipHostMalloc((void **)&ctx->pinned, sizeof(int) * (size_t)nInst, hipHostMallocDefault));
        ctx->pinnedCount = nInst;
    }
    OD_HIP(hipMemcpyAsync(ctx->pinned, doneDev, sizeof(int) * (size_t)nInst, hipMemcpyDeviceToHost, ctx->stream));
    OD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_flush(ctx);
    int n = 0;
    for (int i = 0; i < nInst; i++) n += ctx->pinned[i] ? 0 : 1;
    return n;
}

// Builds the preconditioner from A (Preconditioner::Preprocess, cgSolver.cpp:23).  The FP64
// lower copy of the IC(0) factor is written when the FP64 sweeps will run (fp64Sweeps).
// Threshold IC or SAINV LDL^T of A: the values come to the host, the factor and its sweep
// schedule are built there, and the sweep arrays go back in one upload.
static void build_scalar_precond(od_solver *s, const od_matrix *A) {
    od_mesh *mesh = s->mesh;
    hipStream_t st = mesh->ctx->stream;
    ScalarPrecond &P = *s->scalar;
    const Topology &t = mesh->topo;
    P.values.resize((size_t)t.nValues);
    OD_HIP(hipMemcpyAsync(P.values.data(), A->vals, sizeof(double) * (size_t)t.nValues, hipMemcpyDeviceToHost, st));
    OD_HIP(hipStreamSynchronize(st));
    if (s->precond == OD_PRE_ICT) thresholdIcFactor(t, P.values.data(), s->param, s->param2, P.factor);
    else sainvLdltFactor(t, P.values.data(), s->param, s->param2, P.factor);
    buildScalarSweeps(P.factor, s->precond == OD_PRE_LDLT, P.host);
    const ScalarSweepsHost &h = P.host;
    // one buffer: the int arrays, then (8-byte aligned) the double arrays
    const std::vector<const std::vector<int> *> ints = {&h.fPtr, &h.fCol, &h.bPtr, &h.bRow, &h.fwd.levelPtr,
                                                        &h.fwd.levelRows, &h.bwd.levelPtr, &h.bwd.levelRows};
    const std::vector<const std::vector<double> *> dbls = {&h.fVal, &h.bVal, &h.diag};
    size_t nInt = 0, nDbl = 0;
    for (auto *v : ints) nInt += v->size();
    for (auto *v : dbls) nDbl += v->size();
    const size_t intBytes = (sizeof(int) * nInt + 255) & ~(size_t)255;
    const size_t bytes = intBytes + sizeof(double) * nDbl + 256;
    if (bytes > P.cap) {
        OD_HIP(hipStreamSynchronize(st));
        if (P.buf) OD_HIP(hipFree(P.buf));
        P.buf = nullptr;
        P.cap = 0;
        OD_HIP(hipMalloc(&P.buf, bytes));
        P.cap = bytes;
    }
    std::vector<char> staging(bytes, 0);
    std::vector<const void *> devPtr;
    size_t off = 0;
    for (auto *v : ints) {
        std::memcpy(staging.data() + off, v->data(), sizeof(int) * v->size());
        devPtr.push_back((const char *)P.buf + off);
        off += sizeof(int) * v->size();
    }
    off = intBytes;
    for (auto *v : dbls) {
        std::memcpy(staging.data() + off, v->data(), sizeof(double) * v->size());
        devPtr.push_back((const char *)P.buf + off);
        off += sizeof(double) * v->size();
    }
    OD_HIP(hipMemcpyAsync(P.buf, staging.data(), bytes, hipMemcpyHostToDevice, st));
    OD_HIP(hipStreamSynchronize(st));   // the staging buffer is freed on return
    ScalarSweepsDev &d = P.dev;
    d.n = h.n;
    d.unit = h.unit;
    d.fPtr = (const int *)devPtr[0]; d.fCol = (const int *)devPtr[1];
    d.bPtr = (const int *)devPtr[2]; d.bRow = (const int *)devPtr[3];
    d.fLevelPtr = (const int *)devPtr[4]; d.fLevelRows = (const int *)devPtr[5];
    d.bLevelPtr = (const int *)devPtr[6]; d.bLevelRows = (const int *)devPtr[7];
    d.fVal = (const double *)devPtr[8]; d.bVal = (const double *)devPtr[9]; d.diag = (const double *)devPtr[10];
    d.fPlan = &h.fwd;
    d.bPlan = &h.bwd;
}

static void build_precond(od_solver *s, const od_matrix *A, const uint8_t *skip, bool fp64Sweeps = true) {
    od_mesh *mesh = s->mesh;
    const DevMesh &m = mesh->dm;
    hipStream_t st = mesh->ctx->stream;
    if (s->scalar) {
        build_scalar_precond(s, A);
        return;
    }
    if (s->precond == OD_PRE_JACOBI) {
        launch_jacobi_diag<double>(m, st, A->vals, s->invDiag);
        if (s->invDiag32) launch_jacobi_diag<float>(m, st, A->vals, s->invDiag32);
    } else if (s->precond == OD_PRE_IC0) {
        if (fp64Sweeps) {
            alloc_fp64_sweeps(s);
            launch_ic0_factor<double>(m, st, mesh->topo.levelPtr, A->vals, s->U, s->U64, s->Ul64, s->param, skip);
        } else {
            launch_ic0_factor<float>(m, st, mesh->topo.levelPtr, A->vals, s->U, s->U32, s->Ul32, s->param, skip);
        }
    }
}

// PCG on the chunked instances: x in/out, b fixed.  preDone marks instances to skip.
// FP32 solves fuse the direction update into the product: p holds two direction buffers of
// ndof each, iteration k reads buffer k & 1 and its product writes the new direction
// z + beta p into the other one.  FP64 solves update p in place after the dot product (the
// fused FP64 descriptor product needs more registers than two workgroups per CU allow, and
// the fused grouped FP64 product measured 74 us against 63 us plus 6.6 us for this pass).
template <typename T>
static void run_pcg(od_solver *s, const T *vals, const T *Uf, const T *Ul, const T *invDiag, const T *b,
                    T *x, T *r, T *z, T *p, T *q, double tol, int maxIt, const uint8_t *preDone, InstScalars *sc,
                    int *doneFlags, bool xZero, int expect, const InstScalars *outer = nullptr, int fixedIters = 0) {
    od_mesh *mesh = s->mesh;
    const DevMesh &m = mesh->dm;
    hipStream_t st = mesh->ctx->stream;
    PcgWork w{s->partials, sc, doneFlags, outer};
    if (sizeof(T) == 4) w.slotVals = m.spmvL32 == -1 ? (const void *)s->gv32 : (const void *)s->sv32;
    else if (m.spmvL64 == -1) w.slotVals = s->gv64;
    const size_t n = (size_t)mesh->topo.ndof;
    if (xZero && fixedIters) {   // under step-graph capture: a kernel, not a memset node
        const size_t words = sizeof(T) * n / 4;
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, (uint32_t *)q, words);
        OD_HIP(hipGetLastError());
    } else if (xZero) {
        OD_HIP(hipMemsetAsync(q, 0, sizeof(T) * n, st));
    } else {
        launch_spmv<T>(m, st, vals, x, q, 0, nullptr, (const T *)w.slotVals);
    }
    pcg_init_residual<T>(m, st, b, q, r, w);
    const bool ic = s->precond == OD_PRE_IC0;
    const bool sc64 = s->scalar != nullptr;   // threshold IC / LDL^T (FP64 solvers only)
    auto scalar_apply = [&](const T *rr, T *zz) {
        if constexpr (sizeof(T) == 8) launch_scalar_apply(s->scalar->dev, st, rr, zz);
    };
    if (ic) launch_ic0_apply<T>(m, st, mesh->topo.levelPtr, Uf, Ul, r, z, nullptr, preDone);
    else if (sc64) scalar_apply(r, z);
    else pcg_jacobi<T>(m, st, invDiag, r, z, w);
    pcg_dot_rz_init<T>(m, st, r, z, p, w, tol, maxIt, preDone);
    // Iterations are launched in batches between convergence polls (a poll synchronises the
    // stream).  Warm-started solves of consecutive steps need nearly the same count, so the
    // first batch is sized from the same solve of the previous step (`expect`); afterwards
    // the batches stay small.
    constexpr bool fuseDir = sizeof(T) == 4;
    auto iteration = [&](int k) {
        hipEvent_t e0, e1;
        prof_pair(mesh->ctx, &e0, &e1);
        const T *pOld = fuseDir ? p + (size_t)(k & 1) * n : p;
        T *pNew = fuseDir ? p + (size_t)((k + 1) & 1) * n : p;
        pcg_alpha<T>(m, st, vals, pOld, fuseDir ? z : nullptr, pNew, q, w, e0, e1);
        pcg_alpha_finalize(m, st, w);
        if (ic) {
            pcg_update_xr_ic0<T>(m, st, mesh->topo.levelPtr, pNew, q, x, r, Uf, Ul, z, w, preDone);
            if (ic0_dot_fused(m, mesh->topo.levelPtr)) pcg_finalize_iter(m, st, w);
            else pcg_dot_rz<T>(m, st, r, z, w);
        } else if (sc64) {
            pcg_update_xr<T>(m, st, pNew, q, x, r, nullptr, nullptr, w);   // x, r and r.r only
            scalar_apply(r, z);
            pcg_dot_rz<T>(m, st, r, z, w);
        } else {
            pcg_update_xr<T>(m, st, pNew, q, x, r, invDiag, z, w);
        }
        if (!fuseDir) pcg_update_p<T>(m, st, z, p, w);
    };
    // Graph mode (not while timing launches): a block of kGraphIters iterations is captured
    // once per set of buffers and replayed; iterations past convergence exit at once.
    if (fixedIters) {
        // inside a step graph: exactly this many iterations, each exiting at once past convergence
        for (int j = 0; j < fixedIters; j++) iteration(j);
        return;
    }
    od_solver::IterGraph *g = nullptr;
    const int launchesPerIter = ic ? 2 * m.nLevels + 4 : 5;
    // the scalar preconditioners' sweep schedule changes with every factor: no graph
    if (s->useGraph && !mesh->ctx->prof && !sc64 && kGraphIters * launchesPerIter <= kGraphMaxLaunches) {
        g = &s->graph[sizeof(T) == 8 ? 0 : 1];
        // the buffers and the launch forms (SpMV shapes, sweep form per level) baked into the
        // captured launches
        const std::vector<const void *> key = {vals, Uf, Ul, invDiag, x, r, z, p, q, sc, doneFlags, preDone, w.slotVals,
                                               (const void *)(uintptr_t)mesh->formGen};
        if (!g->exec || g->key != key) {
            if (g->exec) { OD_HIP(hipGraphExecDestroy(g->exec)); g->exec = nullptr; }
            hipGraph_t graph = nullptr;
            OD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (int j = 0; j < kGraphIters; j++) iteration(j);
            } catch (...) {
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            OD_HIP(hipStreamEndCapture(st, &graph));
            const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            OD_HIP(e);
            g->key = key;
            s->captures++;
        }
    }
    // A skipped iteration (all kernels launched, each exiting at once) costs about as much as
    // two polls, so after the first batch the batches stay at 2 iterations.
    int launched = 0;
    int batch = expect > 1 ? expect - 1 : 4;
    while (launched < maxIt) {
        if (count_not_done(mesh->ctx, doneFlags, m.nInst) == 0) break;
        const int k = std::min(batch, maxIt - launched);
        if (g) {
            // whole blocks; the device-side iteration cap (maxIt) still holds
            for (int j = 0; j < k; j += kGraphIters) OD_HIP(hipGraphLaunch(g->exec, st));
            launched += ((k + kGraphIters - 1) / kGraphIters) * kGraphIters;
        } else {
            for (int j = 0; j < k; j++) iteration(launched + j);
            launched += k;
        }
        batch = expect > 1 ? 2 : std::min(batch * 2, 32);
    }
}

static void collect_stats(od_solver *s, const InstScalars *scDev, bool accumulate) {
    int nInst = s->mesh->topo.nInst;
    std::vector<InstScalars> h(nInst);
    OD_HIP(hipMemcpyAsync(h.data(), scDev, sizeof(InstScalars) * nInst, hipMemcpyDeviceToHost, s->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(s->mesh->ctx->stream));
    for (int i = 0; i < nInst; i++) {
        if (accumulate) s->iters[i] += h[i].it;
        else s->iters[i] = h[i].it;
    }
}

static void collect_relres(od_solver *s, const InstScalars *scDev) {
    int nInst = s->mesh->topo.nInst;
    std::vector<InstScalars> h(nInst);
    OD_HIP(hipMemcpyAsync(h.data(), scDev, sizeof(InstScalars) * nInst, hipMemcpyDeviceToHost, s->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(s->mesh->ctx->stream));
    for (int i = 0; i < nInst; i++) s->relres[i] = h[i].bb > 0 ? std::sqrt(h[i].rr / h[i].bb) : std::sqrt(h[i].rr);
}

// Full solve A x = b (x warm start).  skip: per-instance mask of instances left untouched.
// fixed (step graph capture): FP64, {iterations}; mixed, the iterations of each refinement
// sweep.  The launches then need no host decision and the statistics stay on the device.
static void solve(od_solver *s, od_matrix *A, const double *b, double *x, double tol, int maxIt, const uint8_t *skip,
                  const std::vector<int> *fixed = nullptr) {
    od_mesh *mesh = s->mesh;
    const DevMesh &m = mesh->dm;
    hipStream_t st = mesh->ctx->stream;
    const Topology &t = mesh->topo;
    const uint8_t *pre = skip ? skip : s->zeroMask;
    if (s->precision == OD_FP64) {
        build_precond(s, A, pre);
        if (m.spmvL64 == -1) {
            if (!s->gv64) s->gv64 = (double *)s->alloc(sizeof(double) * (size_t)grp_values(m));
            launch_fill_grp<double>(m, st, A->vals, s->gv64);
        }
        run_pcg<double>(s, A->vals, s->U64, s->Ul64, s->invDiag, b, x, s->r, s->z, s->p, s->q, tol, maxIt, pre,
                        s->scIn, s->doneIn, false, s->lastMaxIters, nullptr, fixed ? (*fixed)[0] : 0);
        if (fixed) return;
        collect_stats(s, s->scIn, false);
        collect_relres(s, s->scIn);
        s->outer = 0;
        s->lastMaxIters = *std::max_element(s->iters.begin(), s->iters.end());
        return;
    }
    // ---- mixed: FP32 inner PCG, FP64 residual and correction (iterative refinement)
    require(A->vals32 != nullptr, "mixed precision needs the FP32 copy of A (od_matrix_sync_f32)", OD_ERR_STATE);
    build_precond(s, A, pre, false);
    if (m.spmvL32 == 0) {
        if (!s->sv32) s->sv32 = (float *)s->alloc(sizeof(float) * kSlotValFloats * (size_t)std::max(1, m.nChunks));
        launch_fill_slot_values(m, st, A->vals32, s->sv32);
    } else if (m.spmvL32 == -1) {
        if (!s->gv32) s->gv32 = (float *)s->alloc(sizeof(float) * (size_t)grp_values(m));
        launch_fill_grp<float>(m, st, A->vals32, s->gv32);
    }
    std::fill(s->iters.begin(), s->iters.end(), 0);
    PcgWork wo{s->partials, s->scOut, s->doneOut};
    launch_spmv<double>(m, st, A->vals, x, s->Ax, 0, nullptr);
    launch_residual_f64(m, st, b, s->Ax, s->r, s->r32, wo, tol, pre, 1);
    if (fixed) {
        // the graph's sweeps: each masks converged instances and caps its inner iterations at
        // the plan's count (the plan's total stays within maxIt)
        const int nI = t.nInst;
        const size_t words = t.ndof;
        for (size_t q = 0; q < fixed->size(); q++) {
            hipLaunchKernelGGL(k_done_to_mask, dim3((nI + 255) / 256), dim3(256), 0, st, nI, s->scOut, pre, s->mask);
            hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, (uint32_t *)s->x32,
                               words);
            OD_HIP(hipGetLastError());
            run_pcg<float>(s, A->vals32, s->U32, s->Ul32, s->invDiag32, s->r32, s->x32, s->ri32, s->z32, s->p32,
                           s->q32, s->innerTol, (*fixed)[q], s->mask, s->scIn, s->doneIn, true, 0, s->scOut, (*fixed)[q]);
            hipLaunchKernelGGL(k_record_sweep, dim3((nI + 255) / 256), dim3(256), 0, st, nI, s->scIn,
                               s->sweepRec + q * (size_t)nI);
            OD_HIP(hipGetLastError());
            launch_add_correction(m, st, s->x32, x, s->scOut);
            launch_spmv<double>(m, st, A->vals, x, s->Ax, 0, nullptr);
            launch_residual_f64(m, st, b, s->Ax, s->r, s->r32, wo, tol, pre, 0);
        }
        return;
    }
    // The inner iterations of all refinement sweeps count against maxIt together, as the
    // reference's maxIteration caps the whole solve (cgSolver.cpp:31): a sweep gets the
    // iterations the previous ones left.
    int outer = 0;
    for (; outer < s->maxOuter; outer++) {
        const int used = *std::max_element(s->iters.begin(), s->iters.end());
        if (used >= maxIt && outer > 0) break;
        if (count_not_done(mesh->ctx, s->doneOut, t.nInst) == 0) break;
        hipLaunchKernelGGL(k_done_to_mask, dim3((t.nInst + 255) / 256), dim3(256), 0, st, t.nInst, s->scOut, pre, s->mask);
        OD_HIP(hipGetLastError());
        OD_HIP(hipMemsetAsync(s->x32, 0, sizeof(float) * t.ndof, st));
        run_pcg<float>(s, A->vals32, s->U32, s->Ul32, s->invDiag32, s->r32, s->x32, s->ri32, s->z32, s->p32,
                       s->q32, s->innerTol, maxIt - used, s->mask, s->scIn, s->doneIn, true, s->sweepIters[outer],
                       s->scOut);
        {
            const int before = *std::max_element(s->iters.begin(), s->iters.end());
            collect_stats(s, s->scIn, true);
            const int after = *std::max_element(s->iters.begin(), s->iters.end());
            s->sweepIters[outer] = after - before;
        }
        launch_add_correction(m, st, s->x32, x, s->scOut);
        launch_spmv<double>(m, st, A->vals, x, s->Ax, 0, nullptr);
        launch_residual_f64(m, st, b, s->Ax, s->r, s->r32, wo, tol, pre, 0);
    }
    s->outer = outer;
    collect_relres(s, s->scOut);
}

namespace od {
void set_last_error(const std::string &msg) { g_err = msg; }
}

// =====================================================================================
// C-ABI
// =====================================================================================
extern "C" {

const char *od_last_error(void) { return g_err.c_str(); }
int od_version(void) { return 1; }

int od_ctx_create(int device, od_ctx **out) {
    API_BEGIN
    require(out != nullptr, "out is NULL");
    int count = 0;
    OD_HIP(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "device index out of range");
    OD_HIP(hipSetDevice(device));
    scratch_setup(device);
    od_ctx *c = new od_ctx();
    c->device = device;
    OD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    OD_HIP(hipHostMalloc((void **)&c->pinned, 64 * sizeof(int), hipHostMallocDefault));
    c->pinnedCount = 64;
    *out = c;
    API_END
}

static void ctx_free(od_ctx *ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipHostFree(ctx->pinned);
    delete ctx;
}

// A mesh stops using ctx (destroyed or rebound): the last one frees a closed context.
static void ctx_release(od_ctx *ctx) {
    if (ctx->meshes.fetch_sub(1) == 1 && ctx->closed.load()) ctx_free(ctx);
}

static void mesh_retain(od_mesh *mesh) { mesh->refs.fetch_add(1); }

// Drops one reference; the last one waits for the mesh's stream, frees it and releases its
// context.
static void mesh_release(od_mesh *mesh) {
    if (mesh->refs.fetch_sub(1) != 1) return;
    od_ctx *ctx = mesh->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    delete mesh;
    ctx_release(ctx);
}

int od_ctx_destroy(od_ctx *ctx) {
    API_BEGIN
    if (!ctx) return OD_OK;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->closed.store(true);
    if (ctx->meshes.load() == 0) ctx_free(ctx);   // else the last mesh frees it
    API_END
}

int od_ctx_sync(od_ctx *ctx) {
    API_BEGIN
    require(ctx != nullptr, "ctx is NULL");
    OD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_flush(ctx);
    API_END
}

int od_ctx_profile(od_ctx *ctx, int enable) {
    API_BEGIN
    require(ctx != nullptr, "ctx is NULL");
    OD_HIP(hipStreamSynchronize(ctx->stream));
    if (enable && ctx->ev.empty()) {
        ctx->ev.resize(512);
        for (auto &e : ctx->ev) OD_HIP(hipEventCreate(&e));
    }
    if (ctx->prof) prof_flush(ctx);
    ctx->prof = enable != 0;
    if (enable) { ctx->profMs = 0; ctx->profCount = 0; ctx->evUsed = 0; }
    API_END
}

int od_ctx_profile_read(od_ctx *ctx, double *totalMs, long long *launches) {
    API_BEGIN
    require(ctx && totalMs && launches, "NULL argument");
    OD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_flush(ctx);
    *totalMs = ctx->profMs;
    *launches = ctx->profCount;
    API_END
}

int od_dev_alloc(od_ctx *ctx, long long bytes, void **dev) {
    API_BEGIN
    require(ctx && dev && bytes >= 0, "bad arguments");
    OD_HIP(hipSetDevice(ctx->device));
    OD_HIP(hipMalloc(dev, (size_t)std::max<long long>(bytes, 16)));
    API_END
}
int od_dev_free(od_ctx *ctx, void *dev) {
    API_BEGIN
    require(ctx != nullptr, "ctx is NULL");
    if (dev) OD_HIP(hipFree(dev));
    API_END
}
int od_copy_h2d(od_ctx *ctx, void *dev, const void *host, long long bytes) {
    API_BEGIN
    require(ctx && dev && host && bytes >= 0, "bad arguments");
    OD_HIP(hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    OD_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}
int od_copy_d2h(od_ctx *ctx, void *host, const void *dev, long long bytes) {
    API_BEGIN
    require(ctx && dev && host && bytes >= 0, "bad arguments");
    OD_HIP(hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    OD_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}
int od_memset_d(od_ctx *ctx, void *dev, int byte, long long bytes) {
    API_BEGIN
    require(ctx && dev && bytes >= 0, "bad arguments");
    OD_HIP(hipMemsetAsync(dev, byte, (size_t)bytes, ctx->stream));
    API_END
}

int od_mesh_create(od_ctx *ctx, int nNodes, const double *X, int nTets, const int *tets, int nConstrained,
                   const int *constrained, int nInstances, const int *instNodeStart, const int *instTetStart,
                   od_mesh **out) {
    API_BEGIN
    require(ctx && X && tets && out, "NULL argument");
    require(nNodes > 0 && nTets > 0 && nConstrained >= 0, "empty mesh");
    require(nConstrained == 0 || constrained != nullptr, "constrained list is NULL");
    require(nInstances >= 1, "nInstances must be >= 1");
    require(nInstances == 1 || (instNodeStart && instTetStart), "instance ranges are NULL");
    od_mesh *mesh = new od_mesh();
    try {
        mesh->ctx = ctx;
        Topology &t = mesh->topo;
        t.nNodes = nNodes;
        t.nTets = nTets;
        t.X.assign(X, X + 3 * (size_t)nNodes);
        t.tets.assign(tets, tets + 4 * (size_t)nTets);
        t.constrained.assign(constrained, constrained + nConstrained);
        if (nInstances > 1) {
            t.instNodeStart.assign(instNodeStart, instNodeStart + nInstances + 1);
            t.instTetStart.assign(instTetStart, instTetStart + nInstances + 1);
        }
        std::string err = buildTopology(t, kChunkRows);
        require(err.empty(), err.c_str());
        OD_HIP(hipSetDevice(ctx->device));
        DevMesh &m = mesh->dm;
        m.nNodes = t.nNodes; m.nTets = t.nTets; m.nRows = t.nRows; m.nInst = t.nInst;
        m.nChunks = (int)t.chunkRow0.size(); m.nLevels = t.nLevels;
        m.nBlocks = t.nBlocks; m.nValues = t.nValues; m.nT = (int64_t)t.tRow.size();
        m.X = mesh->upload(t.X);
        m.tets = mesh->upload(t.tets);
        m.nodeRow = mesh->upload(t.nodeRow);
        if (t.partial) {
            m.nodeDof = mesh->upload(t.nodeDof);
            m.vPtr = mesh->upload(t.vPtr);
            m.vSrc = mesh->upload(t.vSrc);
            m.dPtr = mesh->upload(t.dPtr);
            m.dSrc = mesh->upload(t.dSrc);
            for (int64_t v = 0; v < t.nValues; v++)
                for (int k = t.vPtr[v]; k < t.vPtr[v + 1]; k++)
                    if (t.vSrc[k] < 0) mesh->padDiag.push_back(v);
        }
        m.rowPtr = mesh->upload(t.rowPtr);
        m.colIdx = mesh->upload(t.colIdx);
        m.blkOff = mesh->upload(t.blkOff);
        std::vector<int> blkRow(t.nBlocks);
        for (int r = 0; r < t.nRows; r++)
            for (int b = t.rowPtr[r]; b < t.rowPtr[r + 1]; b++) blkRow[b] = r;
        m.blkRow = mesh->upload(blkRow);
        m.valOff = mesh->upload(t.valOff);
        std::vector<int> tOff(t.tBlk.size());
        for (size_t k = 0; k < t.tBlk.size(); k++) tOff[k] = t.blkOff[t.tBlk[k]];
        m.tOff = mesh->upload(tOff);
        m.tPos = mesh->upload(t.tPos);
        m.tPtr = mesh->upload(t.tPtr);
        m.tRow = mesh->upload(t.tRow);
        m.tBlk = mesh->upload(t.tBlk);
        m.pairBlk = mesh->upload(t.pairBlk);
        m.pairSlot = mesh->upload(t.pairSlot);
        m.pairT = mesh->upload(t.pairT);
        m.slotSrc = nullptr; m.rotC = nullptr; m.blkCompact = nullptr; m.gatherSeg = nullptr; m.gatherPerRun = 0;
        if (!t.partial) {
            // the source of each contribution slot, for the compact corotational form
            require(t.nTets < (1 << (31 - kSrcTetShift)), "too many tets for the contribution source encoding",
                    OD_ERR_UNSUPPORTED);
            std::vector<int> src((size_t)std::max<int64_t>(1, t.nPairContrib), 0);
            for (size_t k = 0; k < t.pairSlot.size(); k++)
                if (t.pairSlot[k] >= 0)
                    src[t.pairSlot[k]] = (int)((k / 10) << kSrcTetShift | (k % 10) << 1 | t.pairT[k]);
            m.slotSrc = mesh->upload(src);
        }
        m.blkCStart = mesh->upload(t.blkCStart);
        m.nodeSlot = mesh->upload(t.nodeSlot);
        m.rowCStart = mesh->upload(t.rowCStart);
        m.levelRows = mesh->upload(t.levelRows);
        m.rowLevel = mesh->upload(t.level);
        {
            // IC(0) update pairs (the factor kernel then searches nothing)
            std::vector<int> pp((size_t)t.nBlocks + 1, 0), pr;
            pr.reserve(8 * (size_t)t.nBlocks);
            require(t.nValues < (int64_t)INT32_MAX, "too many values for the IC(0) pair offsets", OD_ERR_UNSUPPORTED);
            for (int I = 0; I < t.nRows; I++)
                for (int b = t.rowPtr[I] + 1; b < t.rowPtr[I + 1]; b++) {
                    const int J = t.colIdx[b];
                    for (int k = t.tPtr[I]; k < t.tPtr[I + 1]; k++) {
                        const int K = t.tRow[k];
                        const int *c0 = t.colIdx.data() + t.tBlk[k] + 1, *c1 = t.colIdx.data() + t.rowPtr[K + 1];
                        const int *f = std::lower_bound(c0, c1, J);
                        if (f == c1 || *f != J) continue;
                        pr.push_back(tOff[k]);
                        pr.push_back(t.blkOff[f - t.colIdx.data()]);
                    }
                    pp[b + 1] = (int)(pr.size() / 2);
                }
            for (int64_t b = 1; b <= t.nBlocks; b++) pp[b] = std::max(pp[b], pp[b - 1]);   // diagonal blocks: empty
            m.icPairPtr = mesh->upload(pp);
            if (pr.empty()) pr.assign(2, 0);
            m.icPairs = (const int2 *)mesh->upload(pr);
        }
        {
            // per-row records of the sweeps, one 16-byte load each
            std::vector<int> fw(4 * (size_t)t.nRows), bw(4 * (size_t)t.nRows);
            for (int r = 0; r < t.nRows; r++) {
                const size_t q = 4 * (size_t)r;
                fw[q] = t.tPtr[r]; fw[q + 1] = t.tPtr[r + 1]; fw[q + 2] = t.valOff[r]; fw[q + 3] = t.rowInst[r];
                bw[q] = t.rowPtr[r]; bw[q + 1] = t.rowPtr[r + 1]; bw[q + 2] = t.valOff[r]; bw[q + 3] = t.rowInst[r];
            }
            m.rowFwd = (const int4 *)mesh->upload(fw);
            m.rowBwd = (const int4 *)mesh->upload(bw);
        }
        {
            // FP32 sweep slots (see DevMesh) and the chunks of each level
            const int nc = (int)t.chunkRow0.size();
            std::vector<int> ch(4 * (size_t)std::max(1, nc), 0), baseF(t.nRows), baseB(t.nRows);
            int64_t oF = 0, oB = 0;
            for (int c = 0; c < nc; c++) {
                int wF = 0, wB = 0;
                for (int I = t.chunkRow0[c]; I < t.chunkRow1[c]; I++) {
                    wF = std::max(wF, t.tPtr[I + 1] - t.tPtr[I]);
                    wB = std::max(wB, t.rowPtr[I + 1] - t.rowPtr[I] - 1);
                }
                require(oF + (int64_t)kChunkRows * wF < INT32_MAX && oB + (int64_t)kChunkRows * wB < INT32_MAX,
                        "too many sweep slots", OD_ERR_UNSUPPORTED);
                ch[4 * (size_t)c] = (int)oF; ch[4 * (size_t)c + 1] = wF;
                ch[4 * (size_t)c + 2] = (int)oB; ch[4 * (size_t)c + 3] = wB;
                for (int I = t.chunkRow0[c]; I < t.chunkRow1[c]; I++) {
                    baseF[I] = (int)oF + (I - t.chunkRow0[c]);
                    baseB[I] = (int)oB + (I - t.chunkRow0[c]);
                }
                oF += (int64_t)kChunkRows * wF;
                oB += (int64_t)kChunkRows * wB;
            }
            m.swNF = oF; m.swNB = oB;
            std::vector<int> colF((size_t)std::max<int64_t>(1, oF), -1), colB((size_t)std::max<int64_t>(1, oB), -1);
            for (int I = 0; I < t.nRows; I++) {
                for (int k = 0; k < t.tPtr[I + 1] - t.tPtr[I]; k++)
                    colF[(size_t)baseF[I] + (size_t)k * kChunkRows] = t.tRow[t.tPtr[I] + k];
                for (int q = 0; q < t.rowPtr[I + 1] - t.rowPtr[I] - 1; q++)
                    colB[(size_t)baseB[I] + (size_t)q * kChunkRows] = t.colIdx[t.rowPtr[I] + 1 + q];
            }
            // per chunk and direction, the halo: the rows its slots' columns name, sorted; the
            // columns become halo slots
            std::vector<int> hcnt(2 * (size_t)std::max(1, nc), 0);
            auto halo = [&](std::vector<int> &col, int oi, int wi, int &H) {
                std::vector<std::vector<int>> lists((size_t)std::max(1, nc));
                H = 1;
                for (int c = 0; c < nc; c++) {
                    auto &L = lists[c];
                    const int64_t o = ch[4 * (size_t)c + oi], e = o + (int64_t)ch[4 * (size_t)c + wi] * kChunkRows;
                    for (int64_t q = o; q < e; q++)
                        if (col[q] >= 0) L.push_back(col[q]);
                    std::sort(L.begin(), L.end());
                    L.erase(std::unique(L.begin(), L.end()), L.end());
                    H = std::max(H, (int)L.size());
                    hcnt[2 * (size_t)c + (oi == 0 ? 0 : 1)] = (int)L.size();
                    for (int64_t q = o; q < e; q++)
                        if (col[q] >= 0) col[q] = (int)(std::lower_bound(L.begin(), L.end(), col[q]) - L.begin());
                }
                std::vector<int> rows((size_t)std::max(1, nc) * H, 0);
                for (int c = 0; c < nc; c++) std::copy(lists[c].begin(), lists[c].end(), rows.begin() + (size_t)c * H);
                return rows;
            };
            int hF = 1, hB = 1;
            const std::vector<int> haloF = halo(colF, 0, 1, hF), haloB = halo(colB, 2, 3, hB);
            m.swHF = hF; m.swHB = hB;
            m.swHaloF = mesh->upload(haloF);
            m.swHaloB = mesh->upload(haloB);
            m.swHaloCnt = (const int2 *)mesh->upload(hcnt);
            m.swChunk = (const int4 *)mesh->upload(ch);
            m.swColF = mesh->upload(colF);
            m.swColB = mesh->upload(colB);
            m.swBaseF = mesh->upload(baseF);
            m.swBaseB = mesh->upload(baseB);
            // per level, the chunks holding its rows, in chunk order
            std::vector<std::pair<int, int>> lcPairs;   // (level, chunk)
            for (int c = 0; c < nc; c++) {
                const size_t first = lcPairs.size();
                for (int I = t.chunkRow0[c]; I < t.chunkRow1[c]; I++) lcPairs.emplace_back(t.level[I], c);
                std::sort(lcPairs.begin() + first, lcPairs.end());
                lcPairs.erase(std::unique(lcPairs.begin() + first, lcPairs.end()), lcPairs.end());
            }
            std::stable_sort(lcPairs.begin(), lcPairs.end(),
                             [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
            // several instances whose chunks mostly lie in one level (at most 1.25 level entries
            // per chunk): the sweeps that finish z form the r.z chunk partials, each chunk's
            // first at its highest level (levels finish in decreasing order)
            m.sweepDotLists = t.nInst > 1 && 4 * (int64_t)lcPairs.size() <= 5 * (int64_t)nc ? 1 : 0;
            {
                std::vector<int> top((size_t)std::max(1, nc), 0);
                for (const auto &pr : lcPairs) top[pr.second] = std::max(top[pr.second], pr.first);
                m.chunkTopLevel = mesh->upload(top);
            }
            mesh->levelChunkPtr.assign(t.nLevels + 1, 0);
            std::vector<int> lc;
            for (const auto &pr : lcPairs) { lc.push_back(pr.second); mesh->levelChunkPtr[pr.first + 1]++; }
            for (int l = 0; l < t.nLevels; l++) mesh->levelChunkPtr[l + 1] += mesh->levelChunkPtr[l];
            if (lc.empty()) lc.push_back(0);
            m.levelChunks = mesh->upload(lc);
            m.levelChunkPtr = mesh->levelChunkPtr.data();
        }
        // lanes per row before tuning: 16 where a level's rows average more than 8 blocks on
        // the swept side, else 8
        mesh->levelLanesFwd.assign(std::max(1, t.nLevels), 8);
        mesh->levelLanesBwd.assign(std::max(1, t.nLevels), 8);
        mesh->levelLanesFwd64.assign(std::max(1, t.nLevels), 8);
        mesh->levelLanesBwd64.assign(std::max(1, t.nLevels), 8);
        for (int l = 0; l < t.nLevels; l++) {
            double nT = 0, nU = 0;
            for (int q = t.levelPtr[l]; q < t.levelPtr[l + 1]; q++) {
                const int I = t.levelRows[q];
                nT += t.tPtr[I + 1] - t.tPtr[I];
                nU += t.rowPtr[I + 1] - t.rowPtr[I] - 1;
            }
            const double cnt = std::max(1, t.levelPtr[l + 1] - t.levelPtr[l]);
            mesh->levelLanesFwd[l] = mesh->levelLanesFwd64[l] = nT / cnt > 8.0 ? 16 : 8;
            mesh->levelLanesBwd[l] = mesh->levelLanesBwd64[l] = nU / cnt > 8.0 ? 16 : 8;
        }
        m.levelLanesFwd = mesh->levelLanesFwd.data();
        m.levelLanesBwd = mesh->levelLanesBwd.data();
        m.levelLanesFwd64 = mesh->levelLanesFwd64.data();
        m.levelLanesBwd64 = mesh->levelLanesBwd64.data();
        m.levelsContiguous = t.levelsContiguous ? 1 : 0;
        m.chunkRow0 = mesh->upload(t.chunkRow0);
        m.chunkRow1 = mesh->upload(t.chunkRow1);
        m.chunkInst = mesh->upload(t.chunkInst);
        m.perRun = t.perRun;
        m.spmvR64 = m.spmvR32 = 1;
        m.spmvL64 = m.spmvL32 = kSpmvLanes;
        {
            // SpMV slots in run order (position p = t.chunkSorted index): kSpmvLanes per row,
            // rows in order (row-loop path for longer rows), and one record per position
            const int nc = m.nChunks;
            std::vector<int> sOff((size_t)nc * kSpmvThreads, 0), sCol((size_t)nc * kSpmvThreads, -1);
            std::vector<int> info(4 * (size_t)nc);
            require(t.nRows <= kSlotColMask, "too many block rows for the SpMV slot encoding", OD_ERR_UNSUPPORTED);
            require(t.nInst <= kRunInstMask, "too many instances for the SpMV run records", OD_ERR_UNSUPPORTED);
            for (int p = 0; p < nc; p++) {
                const int c = t.chunkSorted[p];
                bool lng = false;
                for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) {
                    const size_t base = ((size_t)p * kChunkRows + (r - t.chunkRow0[c])) * kSpmvLanes;
                    int n = 0;
                    auto put = [&](int off, int col) {
                        if (n < kSpmvLanes) { sOff[base + n] = off; sCol[base + n] = col; }
                        n++;
                    };
                    put(t.valOff[r], r | kSlotDiag);
                    for (int b = t.rowPtr[r] + 1; b < t.rowPtr[r + 1]; b++) put(t.blkOff[b], t.colIdx[b]);
                    for (int k = t.tPtr[r]; k < t.tPtr[r + 1]; k++) put(tOff[k], t.tRow[k] | kSlotTrans);
                    if (n > kSpmvLanes) lng = true;
                }
                info[4 * (size_t)p] = c;
                info[4 * (size_t)p + 1] = t.chunkRow0[c];
                info[4 * (size_t)p + 2] = t.chunkRow1[c];
                info[4 * (size_t)p + 3] = t.chunkInst[c] | (lng ? kRunLong : 0);
            }
            m.slotOff = mesh->upload(sOff);
            m.slotCol = mesh->upload(sCol);
            // the columns again in the slot-value order (slot-major per run position)
            std::vector<int> vCol(sCol.size());
            for (size_t q = 0; q < (size_t)nc; q++)
                for (int r = 0; r < kChunkRows; r++)
                    for (int sl = 0; sl < kSpmvLanes; sl++)
                        vCol[q * kSpmvThreads + (size_t)sl * kChunkRows + r] = sCol[q * kSpmvThreads + (size_t)r * kSpmvLanes + sl];
            m.svCol = mesh->upload(vCol);
            m.runInfo = (const int4 *)mesh->upload(info);
        }
        m.spmvR64d = m.spmvR64;
        m.spmvL64d = m.spmvL64;
        if (m.slotSrc) {
            // the compact gather's work: the blocks of each XCD run's chunks, in order, each cut
            // into segments of at most kGatherSeg contributions, a block's segments within one
            // workgroup
            const int nc = (int)t.chunkRow0.size();
            std::vector<std::vector<int>> runs(8);   // 4 ints per segment
            for (int p = 0; p < nc; p++) {
                const int c = t.chunkSorted[p];
                auto &run = runs[std::min(7, p / std::max(1, t.perRun))];
                for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++)
                    for (int b = t.rowPtr[r]; b < t.rowPtr[r + 1]; b++) {
                        const int k0 = t.blkCStart[b], k1 = t.blkCStart[b + 1];
                        const int ns = std::max(1, (k1 - k0 + kGatherSeg - 1) / kGatherSeg);
                        require(ns <= kBlock, "a block has too many contributions for the compact gather",
                                OD_ERR_UNSUPPORTED);
                        const size_t used = (run.size() / 4) % kBlock;
                        if (used + ns > (size_t)kBlock)   // the block's segments start the next workgroup
                            for (size_t q = used; q < (size_t)kBlock; q++) run.insert(run.end(), {-1, 0, 0, -1});
                        for (int i = 0; i < ns; i++)
                            run.insert(run.end(), {b, k0 + i * kGatherSeg, std::min(k1, k0 + (i + 1) * kGatherSeg),
                                                   i == 0 ? ns - 1 : -1});
                    }
            }
            size_t per = 0;
            for (const auto &run : runs) per = std::max(per, (run.size() / 4 + kBlock - 1) / kBlock);
            std::vector<int> seg(4 * 8 * per * kBlock);
            for (size_t q = 0; q < seg.size(); q += 4) { seg[q] = -1; seg[q + 1] = 0; seg[q + 2] = 0; seg[q + 3] = -1; }
            for (int x = 0; x < 8; x++) std::copy(runs[x].begin(), runs[x].end(), seg.begin() + 4 * (size_t)x * per * kBlock);
            m.gatherPerRun = (int)per;
            m.gatherSeg = (const int4 *)mesh->upload(seg);
        }
        {
            // grouped symmetric product: the layout, when the mesh allows it
            GroupLayout g;
            const std::string err = buildGroupLayout(t, kChunkRows, 8, g);
            require(err.empty(), err.c_str(), OD_ERR_UNSUPPORTED);
            m.grpOk64 = m.grpOk32 = 0;
            m.nGroups = m.perRunG = m.grpMaxDeps = m.grpMaxHalo = 0;
            m.grpNu = m.grpNx = 0;
            if (g.ok) {
                m.nGroups = g.nGroups;
                m.perRunG = g.perRun;
                m.grpMaxDeps = g.maxDeps;
                m.grpMaxHalo = g.maxHalo;
                m.grpNu = g.nUp;
                m.grpNx = g.nX;
                m.grpChunks = (const int4 *)mesh->upload(g.runChunks);
                m.grpX = (const int4 *)mesh->upload(g.runX);
                m.grpChunkUp = (const int2 *)mesh->upload(g.chunkUp);
                m.grpRowInfo = mesh->upload(g.rowInfo);
                m.grpUpMeta = mesh->upload(g.upMeta);
                m.grpUpSrc = mesh->upload(g.upSrc);
                m.grpXMeta = (const int2 *)mesh->upload(g.xMeta);
                m.grpXSrc = mesh->upload(g.xSrc);
                m.grpHalo = (const int2 *)mesh->upload(g.grpHalo);
                m.grpHaloRows = mesh->upload(g.haloRows);
                m.grpOk64 = grp_lds_bytes<double>(m, 4) <= lds_max() ? 1 : 0;   // 4 slots per thread: the most LDS
                m.grpOk32 = grp_lds_bytes<float>(m, 4) <= lds_max() ? 1 : 0;
            }
        }
        m.instChunkStart = mesh->upload(t.instChunkStart);
        m.instRowStart = mesh->upload(t.instRowStart);
        m.tetInst = mesh->upload(t.tetInst);
        m.rowInst = mesh->upload(t.rowInst);
        m.kind = (int *)mesh->alloc(sizeof(int) * nTets);
        m.E = (double *)mesh->alloc(sizeof(double) * nTets);
        m.nu = (double *)mesh->alloc(sizeof(double) * nTets);
        m.rho = (double *)mesh->alloc(sizeof(double) * nTets);
        m.thrCorot = 0.0;
        m.thrNeo = 0.1;
        OD_HIP(hipStreamSynchronize(ctx->stream));
        // host copies no longer needed on the per-step path (a partially constrained mesh keeps
        // its 78-entry map: the device expands only the block-level one)
        if (!t.partial) { t.elemMap78.clear(); t.elemMap78.shrink_to_fit(); }
        t.pairSlot.clear(); t.pairSlot.shrink_to_fit();
        t.nodeSlot.clear(); t.nodeSlot.shrink_to_fit();
        t.vPtr.clear(); t.vPtr.shrink_to_fit(); t.vSrc.clear(); t.vSrc.shrink_to_fit();
        t.dPtr.clear(); t.dPtr.shrink_to_fit(); t.dSrc.clear(); t.dSrc.shrink_to_fit();
    } catch (...) {
        delete mesh;
        throw;
    }
    ctx->meshes.fetch_add(1);
    *out = mesh;
    API_END
}

int od_mesh_bind_ctx(od_mesh *mesh, od_ctx *ctx) {
    API_BEGIN
    require(mesh && ctx, "NULL argument");
    require(ctx->device == mesh->ctx->device, "context of another device");
    if (ctx != mesh->ctx) {
        OD_HIP(hipStreamSynchronize(mesh->ctx->stream));
        od_ctx *old = mesh->ctx;
        ctx->meshes.fetch_add(1);
        mesh->ctx = ctx;
        ctx_release(old);
    }
    API_END
}

int od_mesh_destroy(od_mesh *mesh) {
    API_BEGIN
    if (mesh) mesh_release(mesh);   // the mesh lives on while solvers, matrices or integrators use it
    API_END
}

int od_mesh_info(od_mesh *mesh, long long *info) {
    API_BEGIN
    require(mesh && info, "NULL argument");
    const Topology &t = mesh->topo;
    for (int i = 0; i < 16; i++) info[i] = 0;
    info[0] = t.ndofFree; info[1] = t.nRows; info[2] = t.nBlocks; info[3] = t.nValuesRef; info[4] = t.nLevels;
    info[5] = t.nInst; info[6] = t.nTets; info[7] = t.nNodes; info[8] = (long long)t.chunkRow0.size();
    info[9] = t.nPairContrib; info[10] = t.nNodeContrib; info[11] = (long long)mesh->bytes;
    info[12] = (t.levelsContiguous ? 1 : 0) | (mesh->dm.sweepDotLists ? 2 : 0);
    // per level of the first 8, 4 bits each (forward, backward): log2 of the sweep lanes per
    // row, or 8 + log2(K) when the FP32 sweep runs the slot kernel with K slots per thread
    auto lg = [](int v) { return (v >> 5) ? 8 + (v >> 5) : ((v & 31) == 16 ? 4 : ((v & 31) == 4 ? 2 : 3)); };
    for (int l = 0; l < std::min(8, t.nLevels); l++) {
        info[13] |= (long long)lg(mesh->levelLanesFwd[l]) << (4 * l);
        info[14] |= (long long)lg(mesh->levelLanesBwd[l]) << (4 * l);
    }
    info[15] = t.partial ? 1 : 0;
    API_END
}

int od_pattern_export(od_mesh *mesh, int *blockPcol, int *blockColumnOris, int *blockRows, int *elemMap78) {
    API_BEGIN
    require(mesh != nullptr, "mesh is NULL");
    const Topology &t = mesh->topo;
    if (t.partial) {
        if (blockPcol) std::memcpy(blockPcol, t.refPcol.data(), sizeof(int) * t.refPcol.size());
        if (blockColumnOris) std::memcpy(blockColumnOris, t.refOris.data(), sizeof(int) * t.refOris.size());
        if (blockRows) std::memcpy(blockRows, t.refRows.data(), sizeof(int) * t.refRows.size());
        if (elemMap78) std::memcpy(elemMap78, t.elemMap78.data(), sizeof(int) * t.elemMap78.size());
        return OD_OK;
    }
    if (blockPcol) std::memcpy(blockPcol, t.rowPtr.data(), sizeof(int) * (t.nRows + 1));
    if (blockColumnOris) std::memcpy(blockColumnOris, t.valOff.data(), sizeof(int) * (t.nRows + 1));
    if (blockRows)
        for (int64_t b = 0; b < t.nBlocks; b++) blockRows[b] = 3 * t.colIdx[b];
    if (elemMap78) {
        int *dmap = nullptr;
        OD_HIP(hipMalloc(&dmap, sizeof(int) * 78 * (size_t)t.nTets));
        try {
            launch_expand_map(mesh->dm, mesh->ctx->stream, dmap);
            OD_HIP(hipMemcpyAsync(elemMap78, dmap, sizeof(int) * 78 * (size_t)t.nTets, hipMemcpyDeviceToHost, mesh->ctx->stream));
            OD_HIP(hipStreamSynchronize(mesh->ctx->stream));
        } catch (...) {
            (void)hipFree(dmap);
            throw;
        }
        OD_HIP(hipFree(dmap));
    }
    API_END
}

int od_mesh_set_plastic(od_mesh *mesh, const double *yield, const double *creep, const double *maxStrain) {
    API_BEGIN
    require(mesh && yield && creep && maxStrain, "NULL argument");
    const int n = mesh->topo.nTets;
    for (int i = 0; i < n; i++)
        require(yield[i] >= 0 && creep[i] >= 0 && maxStrain[i] >= 0, "plastic parameters must be non-negative");
    hipStream_t st = mesh->ctx->stream;
    DevMesh &m = mesh->dm;
    if (!m.plStrain) {
        m.plYield = (double *)mesh->alloc(sizeof(double) * n);
        m.plCreep = (double *)mesh->alloc(sizeof(double) * n);
        m.plMax = (double *)mesh->alloc(sizeof(double) * n);
        m.plStrain = (double *)mesh->alloc(sizeof(double) * 9 * n);
    }
    OD_HIP(hipMemcpyAsync(m.plYield, yield, sizeof(double) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemcpyAsync(m.plCreep, creep, sizeof(double) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemcpyAsync(m.plMax, maxStrain, sizeof(double) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemsetAsync(m.plStrain, 0, sizeof(double) * 9 * n, st));
    OD_HIP(hipStreamSynchronize(st));
    API_END
}

int od_mesh_plastic_strain(od_mesh *mesh, double *get, const double *set) {
    API_BEGIN
    require(mesh != nullptr, "NULL argument");
    require(mesh->dm.plStrain != nullptr, "the mesh has no plastic tets", OD_ERR_STATE);
    const size_t b = sizeof(double) * 9 * (size_t)mesh->topo.nTets;
    hipStream_t st = mesh->ctx->stream;
    if (get) OD_HIP(hipMemcpyAsync(get, mesh->dm.plStrain, b, hipMemcpyDeviceToHost, st));
    if (set) OD_HIP(hipMemcpyAsync(mesh->dm.plStrain, set, b, hipMemcpyHostToDevice, st));
    OD_HIP(hipStreamSynchronize(st));
    API_END
}

int od_mesh_spmv_variant(od_mesh *mesh, int r64, int lanes64, int r32, int lanes32) {
    API_BEGIN
    require(mesh != nullptr, "NULL argument");
    auto okR = [](int r) { return r == 1 || r == 2 || r == 4; };
    auto okL = [](int l) { return l == 4 || l == 8 || l == 16; };
    auto okRK = [&](int r, int l) { return l == -1 ? (r == 4 || r == 8) : okR(r); };
    require(okRK(r64, lanes64) && okRK(r32, lanes32),
            "chunks per SpMV workgroup must be 1, 2 or 4 (slots per thread 4 or 8 for the grouped form)");
    require((okL(lanes64) || lanes64 == -1) && (okL(lanes32) || lanes32 == 0 || lanes32 == -1),
            "lanes per SpMV block row must be 4, 8 or 16 (or 0 for the FP32 slot-value form, -1 for the grouped form)");
    require(lanes64 != -1 || mesh->dm.grpOk64, "the grouped FP64 product is not available on this mesh",
            OD_ERR_UNSUPPORTED);
    require(lanes32 != -1 || mesh->dm.grpOk32, "the grouped FP32 product is not available on this mesh",
            OD_ERR_UNSUPPORTED);
    if (lanes64 != -1) { mesh->dm.spmvR64d = r64; mesh->dm.spmvL64d = lanes64; }
    mesh->dm.spmvR64 = r64;
    mesh->dm.spmvL64 = lanes64;
    mesh->dm.spmvR32 = r32;
    mesh->dm.spmvL32 = lanes32;
    mesh->spmvTuned = true;
    mesh->formGen++;
    API_END
}

int od_mesh_sweep_form(od_mesh *mesh, int lanes, int slots) {
    API_BEGIN
    require(mesh != nullptr, "NULL argument");
    require(lanes == 4 || lanes == 8 || lanes == 16, "lanes per sweep row must be 4, 8 or 16");
    require(slots == 0 || slots == 2 || slots == 4 || slots == 8 || slots == 16,
            "slots per sweep thread must be 0 (lane-per-row kernel), 2, 4, 8 or 16");
    const int sel = slots == 0 ? 0 : (slots == 2 ? 1 : (slots == 4 ? 2 : (slots == 8 ? 3 : 4)));
    const uint8_t code = (uint8_t)(lanes | (sel << 5));
    std::fill(mesh->levelLanesFwd.begin(), mesh->levelLanesFwd.end(), code);
    std::fill(mesh->levelLanesBwd.begin(), mesh->levelLanesBwd.end(), code);
    std::fill(mesh->levelLanesFwd64.begin(), mesh->levelLanesFwd64.end(), code);
    std::fill(mesh->levelLanesBwd64.begin(), mesh->levelLanesBwd64.end(), code);
    mesh->spmvTuned = true;
    mesh->formGen++;
    API_END
}

int od_mesh_autotune(od_mesh *mesh, double *ms, int *chosen) {
    API_BEGIN
    require(mesh != nullptr, "NULL argument");
    mesh->spmvTuned = false;
    autotune_spmv(mesh);
    if (ms)
        for (int p = 0; p < 2; p++)
            for (int lv = 0; lv < 3; lv++)
                for (int v = 0; v < 3; v++) ms[9 * p + 3 * lv + v] = mesh->tuneMs[p][lv][v];
    if (ms)
        for (int v = 0; v < 3; v++) ms[18 + v] = mesh->tuneSvMs[v];
    if (ms)
        for (int p = 0; p < 2; p++)
            for (int kv = 0; kv < 2; kv++) ms[21 + 2 * p + kv] = mesh->tuneGrpMs[p][kv];
    if (chosen) {
        chosen[0] = mesh->dm.spmvR64; chosen[1] = mesh->dm.spmvL64;
        chosen[2] = mesh->dm.spmvR32; chosen[3] = mesh->dm.spmvL32;
    }
    API_END
}

int od_mesh_rows(od_mesh *mesh, int *instRowStart, int *level) {
    API_BEGIN
    require(mesh != nullptr, "mesh is NULL");
    const Topology &t = mesh->topo;
    if (instRowStart) std::memcpy(instRowStart, t.instRowStart.data(), sizeof(int) * (t.nInst + 1));
    if (level) std::memcpy(level, t.level.data(), sizeof(int) * t.nRows);
    API_END
}

int od_mesh_set_material(od_mesh *mesh, const int *kind, const double *E, const double *nu, const double *rho,
                         double corotThreshold, double neoThreshold) {
    API_BEGIN
    require(mesh && kind && E && nu && rho, "NULL argument");
    const int n = mesh->topo.nTets;
    for (int i = 0; i < n; i++) {
        require(kind[i] == OD_MAT_COROTATIONAL || kind[i] == OD_MAT_NEOHOOKEAN || kind[i] == OD_MAT_COROT_PLASTIC,
                "unknown material kind", OD_ERR_UNSUPPORTED);
        require(E[i] > 0 && nu[i] > -1.0 && nu[i] < 0.5 && rho[i] >= 0, "material parameters out of range");
    }
    hipStream_t st = mesh->ctx->stream;
    OD_HIP(hipMemcpyAsync(mesh->dm.kind, kind, sizeof(int) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemcpyAsync(mesh->dm.E, E, sizeof(double) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemcpyAsync(mesh->dm.nu, nu, sizeof(double) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemcpyAsync(mesh->dm.rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
    OD_HIP(hipStreamSynchronize(st));
    mesh->dm.thrCorot = corotThreshold;
    mesh->dm.thrNeo = neoThreshold;
    mesh->dm.hasKind[0] = mesh->dm.hasKind[1] = mesh->dm.hasKind[2] = 0;
    for (int i = 0; i < n; i++) mesh->dm.hasKind[kind[i]] = 1;
    if (mesh->dm.hasKind[0] && mesh->dm.slotSrc && !mesh->dm.rotC) {   // compact corotational form
        mesh->dm.rotC = (double *)mesh->alloc(sizeof(double) * 16 * (size_t)n);
        mesh->dm.blkCompact = (uint8_t *)mesh->alloc(std::max<int64_t>(1, mesh->dm.nBlocks));
    }
    if (mesh->dm.slotSrc) {
        launch_mark_stored(mesh->dm, st, mesh->topo.nPairContrib);
        OD_HIP(hipStreamSynchronize(st));
    }
    if (mesh->dm.hasKind[2] && !mesh->dm.plStrain) {
        // plastic tets before od_mesh_set_plastic: never yielding (the corotational response)
        std::vector<double> inf(n, HUGE_VAL), zero(n, 0.0);
        mesh->dm.plYield = mesh->upload(inf);
        mesh->dm.plCreep = mesh->upload(zero);
        mesh->dm.plMax = mesh->upload(inf);
        mesh->dm.plStrain = (double *)mesh->alloc(sizeof(double) * 9 * n);
        OD_HIP(hipMemsetAsync(mesh->dm.plStrain, 0, sizeof(double) * 9 * n, st));
        OD_HIP(hipStreamSynchronize(st));
    }
    mesh->hasMaterial = true;
    API_END
}

int od_body_force(od_mesh *mesh, const double *body, double *f_dev) {
    API_BEGIN
    require(mesh && body && f_dev, "NULL argument");
    mesh->ensureContrib();
    double *db = nullptr;
    size_t bytes = sizeof(double) * 3 * (size_t)mesh->topo.nTets;
    OD_HIP(hipMalloc(&db, bytes));
    OD_HIP(hipMemcpyAsync(db, body, bytes, hipMemcpyHostToDevice, mesh->ctx->stream));
    DofVec<double> f(mesh, 0, f_dev, false, true);
    launch_body_force(mesh->dm, mesh->ctx->stream, db, mesh->nodeC, f.dev);
    f.done();
    OD_HIP(hipStreamSynchronize(mesh->ctx->stream));
    OD_HIP(hipFree(db));
    API_END
}

int od_matrix_create(od_mesh *mesh, od_matrix **out) {
    API_BEGIN
    require(mesh && out, "NULL argument");
    od_matrix *m = new od_matrix();
    m->mesh = mesh;
    mesh_retain(mesh);
    m->retained = true;
    size_t b = sizeof(double) * ((size_t)mesh->topo.nValues + kValuePad);
    if (hipMalloc(&m->vals, b) != hipSuccess) { delete m; throw ApiError(OD_ERR_HIP, "hipMalloc failed for matrix values"); }
    OD_HIP(hipMemsetAsync(m->vals, 0, b, mesh->ctx->stream));
    *out = m;
    API_END
}
int od_matrix_destroy(od_matrix *m) {
    API_BEGIN
    if (m) { (void)hipStreamSynchronize(m->mesh->ctx->stream); delete m; }
    API_END
}
int od_matrix_values(od_matrix *m, double **values) {
    API_BEGIN
    require(m && values, "NULL argument");
    *values = m->vals;
    API_END
}
int od_matrix_upload(od_matrix *m, const double *host) {
    API_BEGIN
    require(m && host, "NULL argument");
    if (m->mesh->topo.partial) {
        std::vector<double> v;
        values_to_kernel(m->mesh, host, v);
        OD_HIP(hipMemcpyAsync(m->vals, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, m->mesh->ctx->stream));
        OD_HIP(hipStreamSynchronize(m->mesh->ctx->stream));
        return OD_OK;
    }
    OD_HIP(hipMemcpyAsync(m->vals, host, sizeof(double) * m->mesh->topo.nValues, hipMemcpyHostToDevice, m->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(m->mesh->ctx->stream));
    API_END
}
int od_matrix_download(od_matrix *m, double *host) {
    API_BEGIN
    require(m && host, "NULL argument");
    if (m->mesh->topo.partial) {
        std::vector<double> v((size_t)m->mesh->topo.nValues);
        OD_HIP(hipMemcpyAsync(v.data(), m->vals, sizeof(double) * v.size(), hipMemcpyDeviceToHost, m->mesh->ctx->stream));
        OD_HIP(hipStreamSynchronize(m->mesh->ctx->stream));
        values_from_kernel(m->mesh, v.data(), host);
        return OD_OK;
    }
    OD_HIP(hipMemcpyAsync(host, m->vals, sizeof(double) * m->mesh->topo.nValues, hipMemcpyDeviceToHost, m->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(m->mesh->ctx->stream));
    API_END
}
int od_matrix_zero(od_matrix *m) {
    API_BEGIN
    require(m != nullptr, "NULL argument");
    OD_HIP(hipMemsetAsync(m->vals, 0, sizeof(double) * m->mesh->topo.nValues, m->mesh->ctx->stream));
    API_END
}

__global__ static void k_axpby(double *y, double a, double b, const double *x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * y[i] + (x ? b * x[i] : 0.0);
}
int od_matrix_axpby(od_matrix *m, double a, double b, const od_matrix *o) {
    API_BEGIN
    require(m != nullptr, "NULL argument");
    require(!o || o->mesh == m->mesh, "matrices of different meshes");
    int64_t n = m->mesh->topo.nValues;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->mesh->ctx->stream, m->vals, a, b,
                       o ? o->vals : nullptr, n);
    OD_HIP(hipGetLastError());
    API_END
}

int od_mass_matrix(od_mesh *mesh, od_matrix *M) {
    API_BEGIN
    require(mesh && M && M->mesh == mesh, "bad arguments");
    require(mesh->hasMaterial, "od_mesh_set_material must be called first (density)", OD_ERR_STATE);
    mesh->ensureContrib();
    ElemParams P{1, 0, 0, 0, 0, 0, 0};
    launch_element(mesh->dm, mesh->ctx->stream, nullptr, nullptr, nullptr, P, mesh->blkC, mesh->nodeC, nullptr);
    launch_gather_blocks(mesh->dm, mesh->ctx->stream, mesh->blkC, 1.0, nullptr, 0.0, M->vals, 1, nullptr, nullptr, false);
    API_END
}

int od_assemble(od_mesh *mesh, const double *u_dev, od_matrix *K, double *fint_dev) {
    API_BEGIN
    require(mesh && K && K->mesh == mesh, "bad arguments");
    require(mesh->hasMaterial, "od_mesh_set_material must be called first", OD_ERR_STATE);
    mesh->ensureContrib();
    ElemParams P{0, fint_dev ? 1 : 0, 0, 0, 0, 0, 0};
    DofVec<double> u(mesh, 0, u_dev, true, false), f(mesh, 1, fint_dev, true, true);
    launch_element(mesh->dm, mesh->ctx->stream, u.dev, nullptr, nullptr, P, mesh->blkC, mesh->nodeC, nullptr);
    launch_gather_blocks(mesh->dm, mesh->ctx->stream, mesh->blkC, 1.0, nullptr, 0.0, K->vals, 1, nullptr, nullptr, true);
    if (fint_dev) launch_gather_rhs(mesh->dm, mesh->ctx->stream, mesh->nodeC, nullptr, nullptr, 1.0, f.dev, 1, nullptr);
    f.done();
    API_END
}

int od_spmv(od_matrix *A, const double *x, double *y, int accumulate) {
    API_BEGIN
    require(A && x && y, "NULL argument");
    require(x != y, "x and y must differ (SpMDV asserts src != dest)");
    DofVec<double> xv(A->mesh, 0, x, true, false), yv(A->mesh, 1, y, accumulate != 0, true);
    launch_spmv<double>(A->mesh->dm, A->mesh->ctx->stream, A->vals, xv.dev, yv.dev, accumulate, nullptr);
    yv.done();
    API_END
}

int od_matrix_sync_f32(od_matrix *A) {
    API_BEGIN
    require(A != nullptr, "NULL argument");
    int64_t n = A->mesh->topo.nValues;
    if (!A->vals32) {
        OD_HIP(hipMalloc(&A->vals32, sizeof(float) * ((size_t)n + kValuePad)));
        OD_HIP(hipMemsetAsync(A->vals32, 0, sizeof(float) * ((size_t)n + kValuePad), A->mesh->ctx->stream));
    }
    launch_cast_f64_f32(A->mesh->ctx->stream, A->vals, A->vals32, n);
    API_END
}

int od_spmv_f32(od_matrix *A, const float *x, float *y, int accumulate) {
    API_BEGIN
    require(A && x && y && x != y, "bad arguments");
    require(A->vals32 != nullptr, "call od_matrix_sync_f32 first", OD_ERR_STATE);
    DofVec<float> xv(A->mesh, 0, x, true, false), yv(A->mesh, 1, y, accumulate != 0, true);
    launch_spmv<float>(A->mesh->dm, A->mesh->ctx->stream, A->vals32, xv.dev, yv.dev, accumulate, nullptr);
    yv.done();
    API_END
}

int od_spmv_timed(od_matrix *A, const void *x, void *y, int iters, int fp32, double *avg_ms) {
    API_BEGIN
    require(A && x && y && avg_ms && iters > 0 && x != y, "bad arguments");
    require(!fp32 || A->vals32, "call od_matrix_sync_f32 first", OD_ERR_STATE);
    hipStream_t st = A->mesh->ctx->stream;
    // padded copies (partially constrained nodes) are made outside the timed launches
    DofVec<float> x32(A->mesh, 0, fp32 ? (const float *)x : nullptr, true, false), y32(A->mesh, 1, fp32 ? (float *)y : nullptr, false, true);
    DofVec<double> x64(A->mesh, 0, fp32 ? nullptr : (const double *)x, true, false),
        y64(A->mesh, 1, fp32 ? nullptr : (double *)y, false, true);
    hipEvent_t e0, e1;
    OD_HIP(hipEventCreate(&e0));
    OD_HIP(hipEventCreate(&e1));
    OD_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < iters; k++) {
        if (fp32) launch_spmv<float>(A->mesh->dm, st, A->vals32, x32.dev, y32.dev, 0, nullptr);
        else launch_spmv<double>(A->mesh->dm, st, A->vals, x64.dev, y64.dev, 0, nullptr);
    }
    OD_HIP(hipEventRecord(e1, st));
    y32.done();
    y64.done();
    OD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    OD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / iters;
    API_END
}

int od_integrator_matrix(od_integrator *it, od_matrix **A) {
    API_BEGIN
    require(it && A, "NULL argument");
    *A = &it->A;
    API_END
}

int od_solver_create(od_mesh *mesh, int precond, double param, int precision, od_solver **out) {
    API_BEGIN
    require(mesh && out, "NULL argument");
    *out = make_solver(mesh, precond, param, precision);
    API_END
}
int od_solver_set_refinement(od_solver *s, double innerTol, int maxOuter) {
    API_BEGIN
    require(s != nullptr, "NULL argument");
    require(innerTol > 0 && innerTol < 1 && maxOuter >= 1, "innerTol must be in (0, 1) and maxOuter >= 1");
    s->innerTol = innerTol;
    s->maxOuter = maxOuter;
    s->sweepIters.assign(maxOuter, 0);
    API_END
}

int od_solver_set_precond_params(od_solver *s, double p1, double p2) {
    API_BEGIN
    require(s != nullptr, "NULL argument");
    require(p1 >= 0 && p2 >= 0, "preconditioner parameters must be non-negative");
    if (s->precond == OD_PRE_IC0) require(p1 > 0, "the IC(0) pivot epsilon must be positive");
    s->param = p1;
    s->param2 = p2;
    API_END
}

int od_solver_set_graph(od_solver *s, int enable) {
    API_BEGIN
    require(s != nullptr, "NULL argument");
    s->useGraph = enable != 0;
    API_END
}

int od_solver_info(od_solver *s, long long *info) {
    API_BEGIN
    require(s && info, "NULL argument");
    info[0] = s->captures;
    info[1] = (long long)s->bytes;
    info[2] = s->outer;
    info[3] = s->iters.empty() ? 0 : *std::max_element(s->iters.begin(), s->iters.end());
    API_END
}

int od_solver_destroy(od_solver *s) {
    API_BEGIN
    if (s) { (void)hipStreamSynchronize(s->mesh->ctx->stream); delete s; }
    API_END
}

int od_pcg_solve(od_solver *s, od_matrix *A, const double *b, double *x, double tol, int maxIter, int *iters,
                 double *relres) {
    API_BEGIN
    require(s && A && b && x && A->mesh == s->mesh, "bad arguments");
    require(tol >= 0 && maxIter >= 0, "tol and maxIter must be non-negative");
    if (s->precision == OD_MIXED) {
        int64_t n = A->mesh->topo.nValues;
        if (!A->vals32) {
            OD_HIP(hipMalloc(&A->vals32, sizeof(float) * ((size_t)n + kValuePad)));
            OD_HIP(hipMemsetAsync(A->vals32, 0, sizeof(float) * ((size_t)n + kValuePad), A->mesh->ctx->stream));
        }
        launch_cast_f64_f32(A->mesh->ctx->stream, A->vals, A->vals32, n);
    }
    DofVec<double> bv(A->mesh, 0, b, true, false), xv(A->mesh, 1, x, true, true);
    solve(s, A, bv.dev, xv.dev, tol, maxIter, nullptr);
    xv.done();
    OD_HIP(hipStreamSynchronize(s->mesh->ctx->stream));
    int nInst = s->mesh->topo.nInst;
    if (iters) std::memcpy(iters, s->iters.data(), sizeof(int) * nInst);
    if (relres) std::memcpy(relres, s->relres.data(), sizeof(double) * nInst);
    API_END
}

__global__ static void k_fill_bytes(uint8_t *p, int n, uint8_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Instances of a system the allocation-free solve accepts: its per-instance scalars come back
// through a fixed host array and the context's pinned mailbox (made for 64 at context creation).
constexpr int kNoAllocMaxInst = 64;

int od_pcg_solve_noalloc(od_matrix *A, int precond, const double *b, double *x, double tol, int maxIter, int *iters,
                         double *relres) {
    API_BEGIN
    require(A && b && x, "NULL argument");
    require(!A->mesh->topo.partial, "the allocation-free solve needs whole-node constraints (no padded copies)",
            OD_ERR_UNSUPPORTED);
    require(tol >= 0 && maxIter >= 0, "tol and maxIter must be non-negative");
    require(precond == OD_PRE_NONE || precond == OD_PRE_JACOBI,
            "the allocation-free solve takes OD_PRE_NONE or OD_PRE_JACOBI", OD_ERR_UNSUPPORTED);
    od_mesh *mesh = A->mesh;
    od_ctx *ctx = mesh->ctx;
    const Topology &t = mesh->topo;
    require(t.nInst <= kNoAllocMaxInst && t.nInst <= ctx->pinnedCount,
            "the allocation-free solve takes at most 64 instances", OD_ERR_UNSUPPORTED);
    unsigned char *base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratchDev[ctx->device].load() : nullptr;
    require(base != nullptr, "the device's scratch was not set up", OD_ERR_STATE);
    // layout of the scratch: r, z, p, q, D^-1 (ndof doubles each), the chunk partials, the
    // per-instance scalars, done flags and skip mask, each 256-byte aligned
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t vb = sizeof(double) * (size_t)t.ndof;
    od_solver s;   // a view of the scratch: holds no device memory of its own
    s.mesh = mesh;
    s.precond = precond;
    s.precision = OD_FP64;
    s.r = (double *)take(vb); s.z = (double *)take(vb); s.p = (double *)take(vb); s.q = (double *)take(vb);
    s.invDiag = (double *)take(vb);
    s.partials = (double *)take(sizeof(double) * 3 * (size_t)t.chunkRow0.size());
    s.scIn = (InstScalars *)take(sizeof(InstScalars) * (size_t)t.nInst);
    s.doneIn = (int *)take(sizeof(int) * (size_t)t.nInst);
    s.zeroMask = (uint8_t *)take((size_t)t.nInst);
    require(off <= kNoAllocScratchBytes,
            "the system is too large for the allocation-free solve's 64 MiB scratch (about 1.6M DOF at most)");
    hipStream_t st = ctx->stream;
    // the skip mask starts at zero (a fill kernel, so the call can be captured in a graph)
    hipLaunchKernelGGL(k_fill_bytes, dim3((t.nInst + 255) / 256), dim3(256), 0, st, s.zeroMask, t.nInst, (uint8_t)0);
    OD_HIP(hipGetLastError());
    build_precond(&s, A, s.zeroMask);
    run_pcg<double>(&s, A->vals, nullptr, nullptr, precond == OD_PRE_JACOBI ? s.invDiag : nullptr, b, x, s.r, s.z, s.p,
                    s.q, tol, maxIter, s.zeroMask, s.scIn, s.doneIn, false, 0);
    InstScalars h[kNoAllocMaxInst];
    OD_HIP(hipMemcpyAsync(h, s.scIn, sizeof(InstScalars) * (size_t)t.nInst, hipMemcpyDeviceToHost, st));
    OD_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < t.nInst; i++) {
        if (iters) iters[i] = h[i].it;
        if (relres) relres[i] = h[i].bb > 0 ? std::sqrt(h[i].rr / h[i].bb) : std::sqrt(h[i].rr);
    }
    API_END
}

// Applies the preconditioner built last (IC(0): the FP64 sweep copies of U; scalar kinds: the
// uploaded sweeps; Jacobi: D^-1).
static void precond_solve(od_solver *s, const double *r, double *z) {
    od_mesh *mesh = s->mesh;
    PcgWork w{s->partials, s->scIn, s->doneIn};
    if (s->precond == OD_PRE_IC0)
        launch_ic0_apply<double>(mesh->dm, mesh->ctx->stream, mesh->topo.levelPtr, s->U64, s->Ul64, r, z, nullptr,
                                 nullptr);
    else if (s->scalar) launch_scalar_apply(s->scalar->dev, mesh->ctx->stream, r, z);
    else pcg_jacobi<double>(mesh->dm, mesh->ctx->stream, s->precond == OD_PRE_JACOBI ? s->invDiag : nullptr, r, z, w);
}

int od_precond_preprocess(od_solver *s, od_matrix *A) {
    API_BEGIN
    require(s && A && A->mesh == s->mesh, "bad arguments");
    build_precond(s, A, s->zeroMask);
    s->preprocessed = true;
    API_END
}

int od_precond_solve(od_solver *s, const double *r, double *z) {
    API_BEGIN
    require(s && r && z, "NULL argument");
    require(r != z, "r and z must differ");
    require(s->preprocessed || s->precond == OD_PRE_NONE, "od_precond_preprocess first", OD_ERR_STATE);
    DofVec<double> rv(s->mesh, 0, r, true, false), zv(s->mesh, 1, z, false, true);
    precond_solve(s, rv.dev, zv.dev);
    zv.done();
    OD_HIP(hipStreamSynchronize(s->mesh->ctx->stream));
    API_END
}

int od_precond_apply(od_solver *s, od_matrix *A, const double *r, double *z) {
    API_BEGIN
    require(s && A && r && z && A->mesh == s->mesh, "bad arguments");
    require(r != z, "r and z must differ");
    od_mesh *mesh = s->mesh;
    build_precond(s, A, s->zeroMask);
    s->preprocessed = true;
    DofVec<double> rv(mesh, 0, r, true, false), zv(mesh, 1, z, false, true);
    precond_solve(s, rv.dev, zv.dev);
    zv.done();
    OD_HIP(hipStreamSynchronize(mesh->ctx->stream));
    API_END
}

int od_ic0_factor(od_solver *s, od_matrix *A, double *L) {
    API_BEGIN
    require(s && A && L && A->mesh == s->mesh, "bad arguments");
    require(s->precond == OD_PRE_IC0, "solver was not created with OD_PRE_IC0", OD_ERR_STATE);
    build_precond(s, A, s->zeroMask);
    if (s->mesh->topo.partial) {
        std::vector<double> v((size_t)s->mesh->topo.nValues);
        OD_HIP(hipMemcpyAsync(v.data(), s->U, sizeof(double) * v.size(), hipMemcpyDeviceToHost, s->mesh->ctx->stream));
        OD_HIP(hipStreamSynchronize(s->mesh->ctx->stream));
        values_from_kernel(s->mesh, v.data(), L);
        return OD_OK;
    }
    OD_HIP(hipMemcpyAsync(L, s->U, sizeof(double) * s->mesh->topo.nValues, hipMemcpyDeviceToHost, s->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(s->mesh->ctx->stream));
    API_END
}

// ---- integrators -------------------------------------------------------------------
static od_integrator *make_integrator(od_mesh *mesh, int kind, int precond, double param, int precision) {
    require(mesh->hasMaterial, "od_mesh_set_material must be called first", OD_ERR_STATE);
    od_integrator *it = new od_integrator();
    try {
        it->mesh = mesh;
        mesh_retain(mesh);
        it->retained = true;
        it->kind = kind;
        it->solver = make_solver(mesh, precond, param, precision);
        mesh->ensureContrib();
        size_t n = (size_t)mesh->topo.ndof, nv = (size_t)mesh->topo.nValues;
        it->M.mesh = mesh;
        it->A.mesh = mesh;
        it->M.vals = (double *)it->alloc(8 * (nv + kValuePad));
        it->A.vals = (double *)it->alloc(8 * (nv + kValuePad));
        if (precision == OD_MIXED) it->A.vals32 = (float *)it->alloc(4 * (nv + kValuePad));
        it->d = (double *)it->alloc(8 * n); it->v = (double *)it->alloc(8 * n);
        it->a = (double *)it->alloc(8 * n); it->aNew = (double *)it->alloc(8 * n);
        it->fext = (double *)it->alloc(8 * n); it->imp = (double *)it->alloc(8 * n);
        it->rhs = (double *)it->alloc(8 * n);
        it->active = (uint8_t *)it->alloc(mesh->topo.nInst);
        it->inactive = (uint8_t *)it->alloc(mesh->topo.nInst);
        it->lastIters.assign(mesh->topo.nInst, 0);
        it->lastRelres.assign(mesh->topo.nInst, 0.0);
        ElemParams P{1, 0, 0, 0, 0, 0, 0};
        launch_element(mesh->dm, mesh->ctx->stream, nullptr, nullptr, nullptr, P, mesh->blkC, mesh->nodeC, nullptr);
        launch_gather_blocks(mesh->dm, mesh->ctx->stream, mesh->blkC, 1.0, nullptr, 0.0, it->M.vals, 0, nullptr, nullptr,
                             false);
        OD_HIP(hipStreamSynchronize(mesh->ctx->stream));
    } catch (...) {
        delete it;
        throw;
    }
    return it;
}

int od_newmark_create(od_mesh *mesh, double beta, double gamma, double dt, double alphaM, double alphaK, int precond,
                      double param, double tol, int maxIter, int precision, int initAccel, od_integrator **out) {
    API_BEGIN
    require(mesh && out, "NULL argument");
    require(dt > 0 && beta > 0 && gamma > 0 && tol >= 0 && maxIter >= 0, "bad integrator parameters");
    od_integrator *it = make_integrator(mesh, 0, precond, param, precision);
    it->beta = beta; it->gamma = gamma; it->h = dt; it->alphaM = alphaM; it->alphaK = alphaK;
    it->tol = tol; it->maxIter = maxIter; it->initAccel = initAccel != 0;
    *out = it;
    API_END
}

int od_backward_euler_create(od_mesh *mesh, double dt, double alphaM, double alphaK, int precond, double param,
                             double tol, int maxIter, od_integrator **out) {
    API_BEGIN
    require(mesh && out, "NULL argument");
    require(dt > 0 && tol >= 0 && maxIter >= 0, "bad integrator parameters");
    od_integrator *it = make_integrator(mesh, 1, precond, param, OD_FP64);
    it->h = dt; it->alphaM = alphaM; it->alphaK = alphaK; it->tol = tol; it->maxIter = maxIter;
    *out = it;
    API_END
}

int od_integrator_set_refinement(od_integrator *it, double innerTol, int maxOuter) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    return od_solver_set_refinement(it->solver, innerTol, maxOuter);
    API_END
}

int od_integrator_device_bytes(od_integrator *it, long long *bytes) {
    API_BEGIN
    require(it && bytes, "NULL argument");
    *bytes = (long long)(it->mesh->bytes + it->bytes + it->solver->bytes);
    API_END
}

int od_integrator_solver(od_integrator *it, od_solver **s) {
    API_BEGIN
    require(it && s, "NULL argument");
    *s = it->solver;
    API_END
}

int od_integrator_set_graph(od_integrator *it, int enable) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    require(enable >= 0 && enable <= 2, "enable must be 0, 1 or 2");
    it->stepGraph = enable == 2;
    return od_solver_set_graph(it->solver, enable);
    API_END
}

int od_integrator_destroy(od_integrator *it) {
    API_BEGIN
    if (it) {
        (void)hipStreamSynchronize(it->mesh->ctx->stream);
        delete it;
    }
    API_END
}

int od_integrator_set_external(od_integrator *it, const double *f) {
    API_BEGIN
    require(it && f, "NULL argument");
    OD_HIP(hipMemcpyAsync(it->fext, f, sizeof(double) * it->mesh->topo.ndofFree, hipMemcpyHostToDevice, it->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    API_END
}
int od_integrator_set_impulse(od_integrator *it, const double *f) {
    API_BEGIN
    require(it && f, "NULL argument");
    OD_HIP(hipMemcpyAsync(it->imp, f, sizeof(double) * it->mesh->topo.ndofFree, hipMemcpyHostToDevice, it->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    it->hasImpulse = true;
    API_END
}
int od_integrator_set_active(od_integrator *it, const unsigned char *active) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    if (!active) { it->hasActive = false; return OD_OK; }
    OD_HIP(hipMemcpyAsync(it->active, active, it->mesh->topo.nInst, hipMemcpyHostToDevice, it->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    it->hasActive = true;
    API_END
}

// One step with eager launches.  actOverride: instances to step (the step graph's remainder),
// in place of the caller's active set.
static void integrator_step_eager(od_integrator *it, const uint8_t *actOverride = nullptr) {
    od_mesh *mesh = it->mesh;
    const DevMesh &m = mesh->dm;
    hipStream_t st = mesh->ctx->stream;
    const uint8_t *act = actOverride ? actOverride : (it->hasActive ? it->active : nullptr);
    od_solver *s = it->solver;
    const uint8_t *skipArg = nullptr;
    if (act) {
        launch_set_done_from_active(m, st, it->inactive, act);
        skipArg = it->inactive;
    }
    const double h = it->h;
    if (it->kind == 0) {
        const double beta = it->beta, gamma = it->gamma, aM = it->alphaM, aK = it->alphaK;
        if (it->steps == 0 && it->initAccel) {
            // consistent initial acceleration: M a0 = f_ext - f_int(d0) - C v0
            ElemParams P0{0, 1, aK, 0.0, 1.0, 0.0, aM};
            launch_element(m, st, it->d, it->v, nullptr, P0, mesh->blkC, mesh->nodeC, act);
            launch_gather_rhs(m, st, mesh->nodeC, it->fext, nullptr, 1.0, it->rhs, 0, act);
            od_solver *ms = make_solver(mesh, OD_PRE_JACOBI, 0.0, OD_FP64);
            try {
                solve(ms, &it->M, it->rhs, it->a, it->tol, it->maxIter, skipArg);
                it->initIters = ms->iters.empty() ? 0 : *std::max_element(ms->iters.begin(), ms->iters.end());
            } catch (...) {
                delete ms;
                throw;
            }
            delete ms;
        }
        ElemParams P{0, 1, h + aK, (0.5 - beta) * h * h + aK * (1.0 - gamma) * h, 1.0, (1.0 - gamma) * h, aM};
        launch_element(m, st, it->d, it->v, it->a, P, mesh->blkC, mesh->nodeC, act);
        const double cK = gamma * h * aK + beta * h * h, cM = 1.0 + gamma * h * aM;
        launch_gather_blocks(m, st, mesh->blkC, cK, it->M.vals, cM, it->A.vals, 0, it->A.vals32, act, true);
        launch_gather_rhs(m, st, mesh->nodeC, it->fext, it->hasImpulse ? it->imp : nullptr, 1.0, it->rhs, 0, act);
        // the solve runs in place on aNew, warm-started from a; a and aNew keep their
        // addresses, so the buffers baked into captured PCG graphs stay valid from step to step
        const size_t vb = sizeof(double) * mesh->topo.ndof;
        OD_HIP(hipMemcpyAsync(it->aNew, it->a, vb, hipMemcpyDeviceToDevice, st));
        solve(s, &it->A, it->rhs, it->aNew, it->tol, it->maxIter, skipArg);
        launch_newmark_correct(m, st, it->d, it->v, it->a, it->aNew, h, beta, gamma, act);
        // inactive instances: their aNew still equals a, so the copy keeps their state
        OD_HIP(hipMemcpyAsync(it->a, it->aNew, vb, hipMemcpyDeviceToDevice, st));
    } else {
        // backward Euler (nonlinearImplicitBackwardEuler.h:82-101); d is the mesh displacement state
        const double aM = it->alphaM, aK = it->alphaK;
        ElemParams P{0, 1, h + aK, 0.0, 1.0, 0.0, aM};
        launch_element(m, st, it->d, it->v, nullptr, P, mesh->blkC, mesh->nodeC, act);
        launch_gather_blocks(m, st, mesh->blkC, h * (h + aK), it->M.vals, 1.0 + h * aM, it->A.vals, 0, nullptr, act,
                             true);
        launch_gather_rhs(m, st, mesh->nodeC, it->fext, it->hasImpulse ? it->imp : nullptr, h, it->rhs, 0, act);
        if (it->hasDamp) {
            // the tangent (h + aK) K + aM M + D: A gains h D, the right-hand side -h D v
            const int64_t nv = mesh->topo.nValues;
            hipLaunchKernelGGL(k_axpby, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, it->A.vals, 1.0, h,
                               (const double *)it->D, nv);
            OD_HIP(hipGetLastError());
            launch_spmv<double>(m, st, it->D, it->v, it->Dv, 0, nullptr);
            launch_axpy_rows(m, st, it->rhs, -h, it->Dv, act);
        }
        // delta_v persists across steps as the warm start (aNew holds it)
        solve(s, &it->A, it->rhs, it->aNew, it->tol, it->maxIter, skipArg);
        launch_be_correct(m, st, it->v, it->aNew, act);
    }
    if (actOverride) return;   // the step graph's caller merges the statistics
    it->hasImpulse = false;
    it->lastIters = s->iters;
    it->lastRelres = s->relres;
    it->lastOuter = s->outer;
    it->steps++;
}


// Graph mode (Newmark, after the first step, IC(0) or Jacobi, no impulse pending, not while
// timing launches): one captured graph per step, replayed.  Iteration plan: the previous
// step's count (per refinement sweep for the mixed form) plus a margin, rounded up to a
// multiple of 4 so that nearby counts share a graph.  Instances the graph cannot finish as
// the eager step would (more iterations needed) are stepped again eagerly from their
// untouched state, the committed ones masked out: every instance gets the eager step's bits.
constexpr int kStepGraphMaxLaunches = 16384;
constexpr size_t kStepGraphCache = 8;
// Margin 4 in steps of 4: on 256 batched boxes a margin of 2 in steps of 2 left instances in
// 9 of 42 steps, whose eager completion cost more than the iterations the tighter plan saves.
constexpr int kStepPlanMargin = 4, kStepPlanQuantum = 4;

static bool step_plan(od_integrator *it, std::vector<int> &plan) {
    od_solver *s = it->solver;
    od_mesh *mesh = it->mesh;
    if (!it->stepGraph || it->kind != 0 || it->steps < 1 || it->hasImpulse || !s->useGraph || mesh->ctx->prof ||
        s->scalar)
        return false;
    if (s->precond != OD_PRE_IC0 && s->precond != OD_PRE_JACOBI && s->precond != OD_PRE_NONE) return false;
    auto up4 = [](int v) { return ((v + kStepPlanMargin + kStepPlanQuantum - 1) / kStepPlanQuantum) * kStepPlanQuantum; };
    plan.clear();
    if (s->precision == OD_FP64) {
        plan.push_back(up4(s->lastMaxIters));
    } else {
        const int sweeps = std::max(1, std::min(s->outer, (int)s->sweepIters.size()));
        if (sweeps > 16) return false;
        for (int q = 0; q < sweeps; q++) plan.push_back(up4(s->sweepIters[q]));
    }
    long long total = 0;
    for (int k : plan) total += k;
    const int perIter = s->precond == OD_PRE_IC0 ? 2 * mesh->dm.nLevels + 4 : 5;
    return total <= it->maxIter && total * perIter <= kStepGraphMaxLaunches;
}

static void newmark_graph_body(od_integrator *it, const std::vector<int> &plan) {
    od_mesh *mesh = it->mesh;
    const DevMesh &m = mesh->dm;
    hipStream_t st = mesh->ctx->stream;
    od_solver *s = it->solver;
    const uint8_t *act = it->hasActive ? it->active : nullptr;
    const uint8_t *skipArg = nullptr;
    if (act) {
        launch_set_done_from_active(m, st, it->inactive, act);
        skipArg = it->inactive;
    }
    const double h = it->h, beta = it->beta, gamma = it->gamma, aM = it->alphaM, aK = it->alphaK;
    ElemParams P{0, 1, h + aK, (0.5 - beta) * h * h + aK * (1.0 - gamma) * h, 1.0, (1.0 - gamma) * h, aM};
    launch_element(m, st, it->d, it->v, it->a, P, mesh->blkC, mesh->nodeC, act);
    const double cK = gamma * h * aK + beta * h * h, cM = 1.0 + gamma * h * aM;
    launch_gather_blocks(m, st, mesh->blkC, cK, it->M.vals, cM, it->A.vals, 0, it->A.vals32, act, true);
    launch_gather_rhs(m, st, mesh->nodeC, it->fext, nullptr, 1.0, it->rhs, 0, act);
    const size_t n = (size_t)mesh->topo.ndof;
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, it->aNew, (const double *)it->a, n);
    OD_HIP(hipGetLastError());
    solve(s, &it->A, it->rhs, it->aNew, it->tol, it->maxIter, skipArg, &plan);
    const int nI = mesh->topo.nInst;
    const bool mixed = s->precision != OD_FP64;
    SweepPlan sp{};
    sp.n = mixed ? (int)plan.size() : 0;
    for (int q = 0; q < sp.n; q++) sp.k[q] = plan[q];
    hipLaunchKernelGGL(k_commit_flags, dim3((nI + 255) / 256), dim3(256), 0, st, nI, mixed ? s->scOut : s->scIn,
                       mixed ? (const int *)s->sweepRec : nullptr, sp, act, it->commit);
    hipLaunchKernelGGL(k_uncommitted, dim3((nI + 255) / 256), dim3(256), 0, st, nI, act, (const uint8_t *)it->commit,
                       it->redo);
    OD_HIP(hipGetLastError());
    launch_newmark_commit(m, st, it->d, it->v, it->a, it->aNew, h, beta, gamma, it->commit);
}

// Returns false when the step is not graph-eligible (the caller steps eagerly).
static bool integrator_step_graph(od_integrator *it) {
    std::vector<int> plan;
    if (!step_plan(it, plan)) return false;
    od_mesh *mesh = it->mesh;
    od_solver *s = it->solver;
    hipStream_t st = mesh->ctx->stream;
    const int nI = mesh->topo.nInst;
    const bool mixed = s->precision != OD_FP64;
    if (!it->commit) {
        it->commit = (uint8_t *)it->alloc(nI);
        it->redo = (uint8_t *)it->alloc(nI);
    }
    if (mixed && !s->sweepRec) s->sweepRec = (int *)s->alloc(sizeof(int) * 16 * (size_t)nI);
    std::vector<int> key = plan;
    key.push_back(mixed ? 1 : 0);
    key.push_back(it->hasActive ? 1 : 0);
    key.push_back((int)(mesh->formGen & 0x7fffffff));
    od_integrator::StepGraph *g = nullptr;
    for (auto &c : it->stepGraphs)
        if (c.key == key) g = &c;
    if (!g) {
        if (it->stepGraphs.size() >= kStepGraphCache) {
            auto lru = std::min_element(it->stepGraphs.begin(), it->stepGraphs.end(),
                                        [](const od_integrator::StepGraph &a, const od_integrator::StepGraph &b) {
                                            return a.lastUse < b.lastUse;
                                        });
            if (lru->exec) OD_HIP(hipGraphExecDestroy(lru->exec));
            it->stepGraphs.erase(lru);
        }
        hipGraph_t graph = nullptr;
        OD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            newmark_graph_body(it, plan);
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        OD_HIP(hipStreamEndCapture(st, &graph));
        od_integrator::StepGraph ng;
        ng.key = key;
        const hipError_t e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        OD_HIP(e);
        it->stepGraphs.push_back(ng);
        g = &it->stepGraphs.back();
        it->graphCaptures++;
    }
    g->lastUse = ++it->stepClock;
    OD_HIP(hipGraphLaunch(g->exec, st));
    it->graphSteps++;
    // statistics and commit flags, one synchronisation
    std::vector<InstScalars> scIn(nI), scOut(nI);
    std::vector<int> rec(mixed ? plan.size() * (size_t)nI : 0);
    std::vector<uint8_t> redo(nI);
    OD_HIP(hipMemcpyAsync(scIn.data(), s->scIn, sizeof(InstScalars) * nI, hipMemcpyDeviceToHost, st));
    if (mixed) {
        OD_HIP(hipMemcpyAsync(scOut.data(), s->scOut, sizeof(InstScalars) * nI, hipMemcpyDeviceToHost, st));
        OD_HIP(hipMemcpyAsync(rec.data(), s->sweepRec, sizeof(int) * rec.size(), hipMemcpyDeviceToHost, st));
    }
    OD_HIP(hipMemcpyAsync(redo.data(), it->redo, nI, hipMemcpyDeviceToHost, st));
    OD_HIP(hipStreamSynchronize(st));
    std::vector<int> iters(nI, 0);
    std::vector<double> relres(nI, 0.0);
    std::vector<int> sweepMax(plan.size(), 0);
    int outer = 0;
    for (int i = 0; i < nI; i++) {
        const InstScalars &f = mixed ? scOut[i] : scIn[i];
        relres[i] = f.bb > 0 ? std::sqrt(f.rr / f.bb) : std::sqrt(f.rr);
        if (!mixed) { iters[i] = scIn[i].it; continue; }
        for (size_t q = 0; q < plan.size(); q++) {
            const int c = rec[q * (size_t)nI + i];
            iters[i] += c;
            sweepMax[q] = std::max(sweepMax[q], c);
            if (c > 0 && !redo[i]) outer = std::max(outer, (int)q + 1);   // redone instances count below
        }
    }
    bool allCommitted = true;
    for (int i = 0; i < nI; i++)
        if (redo[i]) allCommitted = false;
    if (!allCommitted) {
        // the active instances the graph left, stepped again from their untouched state
        integrator_step_eager(it, it->redo);
        for (int i = 0; i < nI; i++)
            if (redo[i]) {
                iters[i] = s->iters[i];
                relres[i] = s->relres[i];
            }
        it->graphFallbacks++;
        outer = std::max(outer, s->outer);
    }
    s->iters = iters;
    s->relres = relres;
    s->lastMaxIters = *std::max_element(iters.begin(), iters.end());
    if (mixed) {
        s->outer = std::max(outer, 1);
        for (size_t q = 0; q < plan.size() && q < s->sweepIters.size(); q++)
            if (allCommitted) s->sweepIters[q] = sweepMax[q];
    } else {
        s->outer = 0;
    }
    it->lastIters = iters;
    it->lastRelres = relres;
    it->lastOuter = s->outer;
    it->steps++;
    return true;
}

static void integrator_step(od_integrator *it) {
    if (!integrator_step_graph(it)) integrator_step_eager(it);
}

int od_integrator_step(od_integrator *it, int nSteps) {
    API_BEGIN
    require(it != nullptr && nSteps >= 0, "bad arguments");
    for (int k = 0; k < nSteps; k++) integrator_step(it);
    OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    if (it->mesh->ctx->prof) prof_flush(it->mesh->ctx);
    API_END
}

int od_integrator_get_state(od_integrator *it, double *d, double *v, double *a) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    size_t b = sizeof(double) * it->mesh->topo.ndofFree;
    hipStream_t st = it->mesh->ctx->stream;
    if (d) OD_HIP(hipMemcpyAsync(d, it->d, b, hipMemcpyDeviceToHost, st));
    if (v) OD_HIP(hipMemcpyAsync(v, it->v, b, hipMemcpyDeviceToHost, st));
    if (a) OD_HIP(hipMemcpyAsync(a, it->kind == 0 ? it->a : it->aNew, b, hipMemcpyDeviceToHost, st));
    OD_HIP(hipStreamSynchronize(st));
    API_END
}

int od_integrator_set_state(od_integrator *it, const double *d, const double *v, const double *a) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    size_t b = sizeof(double) * it->mesh->topo.ndofFree;
    hipStream_t st = it->mesh->ctx->stream;
    if (d) OD_HIP(hipMemcpyAsync(it->d, d, b, hipMemcpyHostToDevice, st));
    if (v) OD_HIP(hipMemcpyAsync(it->v, v, b, hipMemcpyHostToDevice, st));
    if (a) OD_HIP(hipMemcpyAsync(it->kind == 0 ? it->a : it->aNew, a, b, hipMemcpyHostToDevice, st));
    OD_HIP(hipStreamSynchronize(st));
    if (a) it->steps = std::max(it->steps, 1);
    API_END
}

int od_integrator_state_dev(od_integrator *it, double **d, double **v, double **a) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    if (d) *d = it->d;
    if (v) *v = it->v;
    if (a) *a = it->kind == 0 ? it->a : it->aNew;
    API_END
}

int od_integrator_graph_info(od_integrator *it, long long *info) {
    API_BEGIN
    require(it && info, "NULL argument");
    info[0] = it->graphCaptures;
    info[1] = it->graphSteps;
    info[2] = it->graphFallbacks;
    info[3] = (long long)it->stepGraphs.size();
    API_END
}

int od_integrator_stats(od_integrator *it, int *iters, double *relres, int *outer, int *initIters) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    int nInst = it->mesh->topo.nInst;
    if (iters) std::memcpy(iters, it->lastIters.data(), sizeof(int) * nInst);
    if (relres) std::memcpy(relres, it->lastRelres.data(), sizeof(double) * nInst);
    if (outer) *outer = it->lastOuter;
    if (initIters) *initIters = it->initIters;
    API_END
}

int od_backward_euler_update_mesh(od_integrator *it) {
    API_BEGIN
    require(it != nullptr && it->kind == 1, "not a backward-Euler integrator", OD_ERR_STATE);
    launch_axpy_rows(it->mesh->dm, it->mesh->ctx->stream, it->d, it->h, it->v, it->hasActive ? it->active : nullptr);
    OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    API_END
}

// ---- mesh relabelling (host only) ----------------------------------------------------
int od_relabel_gps(int nNodes, int nElems, int nodesPerElem, int *elements, int *newLabels) {
    API_BEGIN
    const char *err = od::relabelGps(nNodes, nElems, nodesPerElem, elements, newLabels);
    require(err == nullptr, err ? err : "");
    API_END
}

// ---- plane collision damping ---------------------------------------------------------
static void plane_damping(od_mesh *mesh, double *D, int nPairs, const double *planes, const int *tets, double coeff,
                          const double *disp, const double *vel) {
    require(nPairs >= 0 && nPairs <= (1 << 27), "number of pairs out of range");
    require(!mesh->topo.partial, "plane damping needs whole-node constraints (its pairs address node blocks)",
            OD_ERR_UNSUPPORTED);
    if (nPairs == 0) return;
    require(planes && tets, "NULL argument");
    for (int k = 0; k < nPairs; k++)
        require(tets[k] >= 0 && tets[k] < mesh->topo.nTets, "collision pair names a tet outside the mesh");
    mesh->ensureDamp(nPairs);
    hipStream_t st = mesh->ctx->stream;
    OD_HIP(hipMemcpyAsync(mesh->damp.planes, planes, sizeof(double) * 4 * (size_t)nPairs, hipMemcpyHostToDevice, st));
    OD_HIP(hipMemcpyAsync(mesh->damp.tets, tets, sizeof(int) * (size_t)nPairs, hipMemcpyHostToDevice, st));
    launch_plane_damping(mesh->dm, st, nPairs, mesh->damp, coeff, disp, vel, D);
    OD_HIP(hipStreamSynchronize(st));
}

int od_plane_damping(od_matrix *D, int nPairs, const double *planes, const int *tets, double coeff, const double *disp,
                     const double *vel) {
    API_BEGIN
    require(D != nullptr, "NULL argument");
    plane_damping(D->mesh, D->vals, nPairs, planes, tets, coeff, disp, vel);
    API_END
}

int od_integrator_add_plane_damping(od_integrator *it, int nPairs, const double *planes, const int *tets,
                                    double coeff) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    require(it->kind == 1, "external damping matrices are a backward-Euler feature (Intergrator reports them unimplemented)",
            OD_ERR_UNSUPPORTED);
    if (!it->D) {
        const size_t nv = (size_t)it->mesh->topo.nValues;
        it->D = (double *)it->alloc(8 * (nv + kValuePad));
        it->Dv = (double *)it->alloc(8 * (size_t)it->mesh->topo.ndof);
    }
    plane_damping(it->mesh, it->D, nPairs, planes, tets, coeff, it->d, it->v);
    it->hasDamp = true;
    API_END
}

int od_integrator_damping_matrix(od_integrator *it, od_matrix **D) {
    API_BEGIN
    require(it && D, "NULL argument");
    require(it->kind == 1, "external damping matrices are a backward-Euler feature (Intergrator reports them unimplemented)",
            OD_ERR_UNSUPPORTED);
    if (!it->D) {
        const size_t nv = (size_t)it->mesh->topo.nValues;
        it->D = (double *)it->alloc(8 * (nv + kValuePad));
        it->Dv = (double *)it->alloc(8 * (size_t)it->mesh->topo.ndof);
    }
    it->Dm.mesh = it->mesh;
    it->Dm.vals = it->D;
    it->hasDamp = true;
    *D = &it->Dm;
    API_END
}

int od_integrator_clear_damping(od_integrator *it) {
    API_BEGIN
    require(it != nullptr, "NULL argument");
    if (it->D) {
        OD_HIP(hipMemsetAsync(it->D, 0, 8 * (size_t)it->mesh->topo.nValues, it->mesh->ctx->stream));
        OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    }
    it->hasDamp = false;
    API_END
}

int od_integrator_damping(od_integrator *it, double *values) {
    API_BEGIN
    require(it && values, "NULL argument");
    const Topology &t = it->mesh->topo;
    const size_t b = 8 * (size_t)t.nValues;
    if (!it->D) {
        std::memset(values, 0, 8 * (size_t)t.nValuesRef);
    } else if (t.partial) {
        std::vector<double> v((size_t)t.nValues);
        OD_HIP(hipMemcpyAsync(v.data(), it->D, b, hipMemcpyDeviceToHost, it->mesh->ctx->stream));
        OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
        values_from_kernel(it->mesh, v.data(), values);
    } else {
        OD_HIP(hipMemcpyAsync(values, it->D, b, hipMemcpyDeviceToHost, it->mesh->ctx->stream));
        OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    }
    API_END
}

int od_integrator_positions(od_integrator *it, float *pos) {
    API_BEGIN
    require(it && pos, "NULL argument");
    const Topology &t = it->mesh->topo;
    std::vector<double> d(t.ndof);
    OD_HIP(hipMemcpyAsync(d.data(), it->d, sizeof(double) * t.ndof, hipMemcpyDeviceToHost, it->mesh->ctx->stream));
    OD_HIP(hipStreamSynchronize(it->mesh->ctx->stream));
    for (int n = 0; n < t.nNodes; n++) {
        int r = t.nodeRow[n];
        for (int a = 0; a < 3; a++) {
            const int o = t.partial ? t.nodeDof[3 * (size_t)n + a] : (r >= 0 ? 3 * r + a : -1);
            pos[3 * n + a] = float(t.X[3 * n + a] + (o >= 0 ? d[o] : 0.0));
        }
    }
    API_END
}

}  // extern "C"
