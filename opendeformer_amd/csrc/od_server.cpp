// Serving loop over the C-ABI: a pool of registered objects (meshes with their
// materials and loads), made device-resident on first use within an HBM budget
// and evicted least-recently-used when the budget needs room.  Requests (advance
// one object by k Newmark steps, with an impulse on its first step) are served
// by nSlots worker slots, each a host thread with its own HIP stream: a request
// that arrives while others run is admitted as soon as a slot and the memory for
// its object are free, and the slots' kernels overlap on the device.
//
// Every object's arithmetic runs on its own mesh and integrator, so its results
// do not depend on what else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/od.h"

namespace od {
void set_last_error(const std::string &msg);   // the text od_last_error() returns (od_capi.hip)
}

namespace {

struct ServerError : std::runtime_error {
    int code;
    ServerError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void check(int status, const char *what) {
    if (status != OD_OK) throw ServerError(status, std::string(what) + ": " + od_last_error());
}

struct Object {
    // host definition
    int nNodes = 0, nTets = 0;
    std::vector<double> X, E, nu, rho, body;
    std::vector<int> tets, constrained, kind;
    double dt = 0.0;
    int ndof = 0;
    // device residency
    od_mesh *mesh = nullptr;
    od_integrator *it = nullptr;
    long long bytes = 0;          // device bytes while resident (estimate before the first load)
    bool measured = false;
    bool busy = false;            // serving a request, being loaded or evicted
    uint64_t lastUse = 0;
    // state kept on the host while not resident
    std::vector<double> d, v, a;
    bool hasState = false;
    long long steps = 0;
    std::vector<int> queue;       // its arrived, not started requests (arrival order)
    size_t queueHead = 0;
};

struct Request {
    int obj = 0, nSteps = 0;
    double arrival = 0.0;
    std::vector<int> impNodes;
    std::vector<double> impForces;
    bool started = false, done = false;
    double start = -1.0, finish = -1.0;
    int iters = 0;
};

}  // namespace

struct od_server {
    int device = 0;
    long long budget = 0;
    int nSlots = 1;
    int precision = OD_MIXED, precond = OD_PRE_IC0;
    double tol = 1e-6;
    int spmvR64 = 0, spmvL64 = 0, spmvR32 = 0, spmvL32 = 0;   // SpMV variant tuned on the first object, reused for the rest
    int fbR64 = 1, fbL64 = 4, fbR32 = 4, fbL32 = 0;           // the best shapes without the grouped form (same bits)
    std::vector<od_ctx *> slotCtx;
    std::vector<Object> objs;
    // run state
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Request> reqs;      // arrival order
    size_t nextArrival = 0;         // first request not yet arrived
    std::vector<int> waiting;       // arrived, not started (arrival order)
    // resident (the admission account) counts the bytes of resident objects, the
    // reservations of objects being loaded and the objects being evicted until their unload
    // has freed them; allocated counts only what is on the device (loaded objects, victims
    // not yet freed), and its maximum is the reported peak
    long long resident = 0, allocated = 0, peak = 0, evictions = 0, loads = 0;
    int nDone = 0;
    uint64_t clock = 0;
    std::chrono::steady_clock::time_point t0;
    std::string failure;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// Device bytes an object needs before it has been loaded once (mesh, contribution slots,
// matrices, factor and sweep copies, grouped values, halos, vectors): measured 1.73-3.03 KB
// per tet on tiled 4-colour boxes from 264 to 500k tets in either precision, the smallest
// boxes the most (structures of fixed size per chunk); 2.5 KB per tet plus 256 KB bounds it.
// The first load replaces it with the measured size.
long long estimate_bytes(const Object &o) { return 2560LL * o.nTets + (256LL << 10); }

// Free DOF of (node, axis), -1 when constrained (core/nodeIndexer.cpp:12-25).
int free_dof(const Object &o, int node, int axis) {
    const int dof = 3 * node + axis;
    auto p = std::lower_bound(o.constrained.begin(), o.constrained.end(), dof);
    if (p != o.constrained.end() && *p == dof) return -1;
    return dof - (int)(p - o.constrained.begin());
}

void unload(od_server *S, Object &o) {
    o.d.resize(o.ndof); o.v.resize(o.ndof); o.a.resize(o.ndof);
    check(od_integrator_get_state(o.it, o.d.data(), o.v.data(), o.a.data()), "od_integrator_get_state");
    o.hasState = true;
    check(od_integrator_destroy(o.it), "od_integrator_destroy");
    check(od_mesh_destroy(o.mesh), "od_mesh_destroy");
    o.it = nullptr;
    o.mesh = nullptr;
    (void)S;
}

// Loads o (its bytes reserved in S->resident) with the launch shapes od_server_run tuned.
// Its allocations start once they fit the budget beside what is on the device: victims of
// other admissions may still be in flight, and their unloads (which never wait) free them.
void load(od_server *S, Object &o, od_ctx *ctx) {
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv.wait(lk, [&] { return S->allocated + o.bytes <= S->budget || !S->failure.empty(); });
        if (!S->failure.empty()) throw ServerError(OD_ERR_STATE, "server run failed elsewhere");
        S->allocated += o.bytes;     // the reservation, replaced by the measured size below
    }
    check(od_mesh_create(ctx, o.nNodes, o.X.data(), o.nTets, o.tets.data(), (int)o.constrained.size(),
                         o.constrained.empty() ? nullptr : o.constrained.data(), 1, nullptr, nullptr, &o.mesh),
          "od_mesh_create");
    check(od_mesh_set_material(o.mesh, o.kind.data(), o.E.data(), o.nu.data(), o.rho.data(), 0.0, 0.1),
          "od_mesh_set_material");
    int sv = od_mesh_spmv_variant(o.mesh, S->spmvR64, S->spmvL64, S->spmvR32, S->spmvL32);
    if (sv == OD_ERR_UNSUPPORTED)   // an object without the grouped form (rows of more than 16 blocks)
        sv = od_mesh_spmv_variant(o.mesh, S->fbR64, S->fbL64, S->fbR32, S->fbL32);
    check(sv, "od_mesh_spmv_variant");
    check(od_newmark_create(o.mesh, 0.25, 0.5, o.dt, 0.0, 0.0, S->precond, 1e-3, S->tol, 10000, S->precision, 1,
                            &o.it),
          "od_newmark_create");
    long long info[16];
    check(od_mesh_info(o.mesh, info), "od_mesh_info");
    o.ndof = (int)info[0];
    std::vector<double> f(o.ndof);
    double *fd = nullptr;
    check(od_dev_alloc(ctx, 8LL * std::max(1, o.ndof), (void **)&fd), "od_dev_alloc");
    const int st = od_body_force(o.mesh, o.body.data(), fd);
    if (st == OD_OK) check(od_copy_d2h(ctx, f.data(), fd, 8LL * o.ndof), "od_copy_d2h");
    od_dev_free(ctx, fd);
    check(st, "od_body_force");
    check(od_integrator_set_external(o.it, f.data()), "od_integrator_set_external");
    if (o.hasState) check(od_integrator_set_state(o.it, o.d.data(), o.v.data(), o.a.data()), "od_integrator_set_state");
    long long bytes = 0;
    check(od_integrator_device_bytes(o.it, &bytes), "od_integrator_device_bytes");
    std::unique_lock<std::mutex> lk(S->mu);
    S->resident += bytes - o.bytes;   // the reservation becomes the measured size
    S->allocated += bytes - o.bytes;
    S->peak = std::max(S->peak, S->allocated);
    o.bytes = bytes;
    o.measured = true;
    S->loads++;
}

// Picks the earliest-arrived request that can start now (the oldest waiting request of its
// object, object idle, memory available); reserves memory for its object, choosing
// least-recently-used idle resident objects to evict.  Returns its index in S->waiting, or
// -1.  Called under the lock.
int pick(od_server *S, double now, std::vector<int> &victims) {
    while (S->nextArrival < S->reqs.size() && S->reqs[S->nextArrival].arrival <= now) {
        const int k = (int)S->nextArrival++;
        S->waiting.push_back(k);
        S->objs[S->reqs[k].obj].queue.push_back(k);
    }
    victims.clear();
    for (size_t w = 0; w < S->waiting.size(); w++) {
        const int k = S->waiting[w];
        Object &o = S->objs[S->reqs[k].obj];
        if (o.busy || o.queue[o.queueHead] != k) continue;
        if (o.mesh) return (int)w;
        const long long need = o.measured ? o.bytes : estimate_bytes(o);
        long long room = S->budget - S->resident;
        if (room < need) {
            std::vector<int> idle;
            for (size_t j = 0; j < S->objs.size(); j++)
                if (S->objs[j].mesh && !S->objs[j].busy) idle.push_back((int)j);
            std::sort(idle.begin(), idle.end(), [&](int a, int b) { return S->objs[a].lastUse < S->objs[b].lastUse; });
            for (int j : idle) {
                if (room >= need) break;
                victims.push_back(j);
                room += S->objs[j].bytes;
            }
            if (room < need) { victims.clear(); continue; }   // not admissible yet: wait for memory
        }
        return (int)w;
    }
    return -1;
}

void worker(od_server *S, int slot) {
    try {
        if (hipSetDevice(S->device) != hipSuccess) throw ServerError(OD_ERR_HIP, "hipSetDevice failed");
        od_ctx *ctx = S->slotCtx[slot];
        std::vector<int> victims;
        for (;;) {
            std::unique_lock<std::mutex> lk(S->mu);
            int k = -1;
            for (;;) {
                if (S->nDone >= (int)S->reqs.size() || !S->failure.empty()) return;
                const double t = S->now();
                const int w = pick(S, t, victims);
                if (w >= 0) {
                    k = S->waiting[w];
                    S->waiting.erase(S->waiting.begin() + w);
                    S->objs[S->reqs[k].obj].queueHead++;
                    break;
                }
                if (S->nextArrival < S->reqs.size())
                    S->cv.wait_for(lk, std::chrono::duration<double>(S->reqs[S->nextArrival].arrival - t));
                else
                    S->cv.wait(lk);
            }
            Request &r = S->reqs[k];
            Object &o = S->objs[r.obj];
            r.started = true;
            r.start = S->now();
            o.busy = true;
            // The victims are unloaded by this worker before its load allocates.  Their bytes
            // stay in the account until each unload has freed them, and the object's
            // reservation enters it now, so no other worker can admit a load into memory
            // that is still held: the account exceeds the budget only by reservations whose
            // allocations wait for these unloads.
            for (int j : victims) S->objs[j].busy = true;
            if (!o.mesh) {
                const long long need = o.measured ? o.bytes : estimate_bytes(o);
                o.bytes = need;              // reservation until the load measures it
                S->resident += need;
            }
            lk.unlock();

            for (int j : victims) {
                Object &v = S->objs[j];
                check(od_mesh_bind_ctx(v.mesh, ctx), "od_mesh_bind_ctx");
                unload(S, v);
                std::lock_guard<std::mutex> g(S->mu);
                S->resident -= v.bytes;
                S->allocated -= v.bytes;
                v.busy = false;
                S->evictions++;
                S->cv.notify_all();
            }
            if (!o.mesh) load(S, o, ctx);
            else check(od_mesh_bind_ctx(o.mesh, ctx), "od_mesh_bind_ctx");
            if (!r.impNodes.empty()) {
                std::vector<double> f(o.ndof, 0.0);
                for (size_t q = 0; q < r.impNodes.size(); q++)
                    for (int a = 0; a < 3; a++) {
                        const int g = free_dof(o, r.impNodes[q], a);
                        if (g >= 0) f[g] += r.impForces[3 * q + a];
                    }
                check(od_integrator_set_impulse(o.it, f.data()), "od_integrator_set_impulse");
            }
            int iters = 0;
            for (int s = 0; s < r.nSteps; s++) {
                check(od_integrator_step(o.it, 1), "od_integrator_step");
                int it = 0;
                check(od_integrator_stats(o.it, &it, nullptr, nullptr, nullptr), "od_integrator_stats");
                iters += it;
            }

            lk.lock();
            r.finish = S->now();
            r.iters = iters;
            r.done = true;
            o.steps += r.nSteps;
            o.busy = false;
            o.lastUse = ++S->clock;
            S->nDone++;
            S->cv.notify_all();
        }
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->failure.empty()) S->failure = e.what();
        S->cv.notify_all();
    }
}

}  // namespace

#define SRV_BEGIN try {
#define SRV_END                                                                   \
    }                                                                             \
    catch (const ServerError &e) { od::set_last_error(e.what()); return e.code; }            \
    catch (const std::exception &e) { od::set_last_error(e.what()); return OD_ERR_ARG; }     \
    return OD_OK;

static void srv_require(bool c, const char *msg, int code = OD_ERR_ARG) {
    if (!c) throw ServerError(code, msg);
}

extern "C" {

int od_server_create(int device, long long hbmBudgetBytes, int nSlots, int precision, int precond, double tol,
                     od_server **out) {
    SRV_BEGIN
    srv_require(out != nullptr, "out is NULL");
    srv_require(hbmBudgetBytes > 0 && nSlots >= 1 && nSlots <= 16 && tol > 0, "bad server parameters");
    srv_require(precision == OD_FP64 || precision == OD_MIXED, "unknown precision", OD_ERR_UNSUPPORTED);
    srv_require(precond == OD_PRE_JACOBI || precond == OD_PRE_IC0, "unknown preconditioner", OD_ERR_UNSUPPORTED);
    od_server *S = new od_server();
    S->device = device;
    S->budget = hbmBudgetBytes;
    S->nSlots = nSlots;
    S->precision = precision;
    S->precond = precond;
    S->tol = tol;
    for (int i = 0; i < nSlots; i++) {
        od_ctx *c = nullptr;
        const int st = od_ctx_create(device, &c);
        if (st != OD_OK) {
            for (od_ctx *p : S->slotCtx) od_ctx_destroy(p);
            delete S;
            check(st, "od_ctx_create");
        }
        S->slotCtx.push_back(c);
    }
    *out = S;
    SRV_END
}

int od_server_destroy(od_server *S) {
    SRV_BEGIN
    if (!S) return OD_OK;
    for (Object &o : S->objs) {
        if (o.it) od_integrator_destroy(o.it);
        if (o.mesh) od_mesh_destroy(o.mesh);
    }
    for (od_ctx *c : S->slotCtx) od_ctx_destroy(c);
    delete S;
    SRV_END
}

int od_server_add_object(od_server *S, int nNodes, const double *X, int nTets, const int *tets, int nConstrained,
                         const int *constrained, const int *kind, const double *E, const double *nu, const double *rho,
                         const double *bodyPerTet, double dt, int *objId) {
    SRV_BEGIN
    srv_require(S && X && tets && kind && E && nu && rho && bodyPerTet && objId, "NULL argument");
    srv_require(nNodes > 0 && nTets > 0 && nConstrained >= 0 && dt > 0, "empty object");
    srv_require(nConstrained == 0 || constrained != nullptr, "constrained list is NULL");
    Object o;
    o.nNodes = nNodes;
    o.nTets = nTets;
    o.X.assign(X, X + 3 * (size_t)nNodes);
    o.tets.assign(tets, tets + 4 * (size_t)nTets);
    o.constrained.assign(constrained, constrained + nConstrained);
    srv_require(std::is_sorted(o.constrained.begin(), o.constrained.end()), "constrained DOFs must be sorted");
    o.kind.assign(kind, kind + nTets);
    o.E.assign(E, E + nTets);
    o.nu.assign(nu, nu + nTets);
    o.rho.assign(rho, rho + nTets);
    o.body.assign(bodyPerTet, bodyPerTet + 3 * (size_t)nTets);
    o.dt = dt;
    o.ndof = 3 * nNodes - nConstrained;
    o.bytes = estimate_bytes(o);
    *objId = (int)S->objs.size();
    S->objs.push_back(std::move(o));
    SRV_END
}

int od_server_run(od_server *S, int nReq, const int *obj, const int *nSteps, const double *arrival,
                  const int *impStart, const int *impNodes, const double *impForces, double *start, double *finish,
                  int *iters, long long *stats) {
    SRV_BEGIN
    srv_require(S && nReq >= 0 && (nReq == 0 || (obj && nSteps && arrival)), "NULL argument");
    S->reqs.assign(nReq, Request());
    for (int k = 0; k < nReq; k++) {
        srv_require(obj[k] >= 0 && obj[k] < (int)S->objs.size(), "unknown object");
        srv_require(nSteps[k] >= 1 && arrival[k] >= 0, "bad request");
        Request &r = S->reqs[k];
        r.obj = obj[k];
        r.nSteps = nSteps[k];
        r.arrival = arrival[k];
        if (impStart) {
            for (int q = impStart[k]; q < impStart[k + 1]; q++) {
                srv_require(impNodes[q] >= 0 && impNodes[q] < S->objs[r.obj].nNodes, "impulse node out of range");
                r.impNodes.push_back(impNodes[q]);
                for (int a = 0; a < 3; a++) r.impForces.push_back(impForces[3 * q + a]);
            }
        }
        const Object &o = S->objs[r.obj];
        srv_require((o.measured ? o.bytes : estimate_bytes(o)) <= S->budget, "an object does not fit the HBM budget");
    }
    // requests in arrival order (stable), so the earliest waiting request is served first
    std::vector<int> order(nReq);
    for (int k = 0; k < nReq; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return S->reqs[a].arrival < S->reqs[b].arrival; });
    std::vector<Request> sorted(nReq);
    for (int k = 0; k < nReq; k++) sorted[k] = S->reqs[order[k]];
    S->reqs.swap(sorted);
    S->nDone = 0;
    S->nextArrival = 0;
    S->waiting.clear();
    for (Object &o : S->objs) { o.queue.clear(); o.queueHead = 0; }
    S->failure.clear();
    if (S->spmvR64 == 0 && nReq > 0) {
        // Launch shapes, tuned once per server on the first request's object before any request
        // is admitted: the tuner's scratch (a 512 MB cache-evicting buffer and buffers of the
        // object's size) is allocated and freed here, outside the residency the budget governs.
        const Object &o = S->objs[S->reqs[0].obj];
        od_mesh *tm = nullptr;
        check(od_mesh_create(S->slotCtx[0], o.nNodes, o.X.data(), o.nTets, o.tets.data(), (int)o.constrained.size(),
                             o.constrained.empty() ? nullptr : o.constrained.data(), 1, nullptr, nullptr, &tm),
              "od_mesh_create");
        int chosen[4];
        double ms[25];
        const int st = od_mesh_autotune(tm, ms, chosen);
        od_mesh_destroy(tm);
        check(st, "od_mesh_autotune");
        S->spmvR64 = chosen[0]; S->spmvL64 = chosen[1];
        S->spmvR32 = chosen[2]; S->spmvL32 = chosen[3];
        // the fastest descriptor / slot-value shapes, for objects without the grouped form
        static const int kR[3] = {1, 2, 4}, kL[3] = {16, 8, 4};
        double b64 = 1e300, b32 = 1e300;
        for (int lv = 0; lv < 3; lv++)
            for (int v = 0; v < 3; v++) {
                if (ms[3 * lv + v] < b64) { b64 = ms[3 * lv + v]; S->fbR64 = kR[v]; S->fbL64 = kL[lv]; }
                if (ms[9 + 3 * lv + v] < b32) { b32 = ms[9 + 3 * lv + v]; S->fbR32 = kR[v]; S->fbL32 = kL[lv]; }
            }
        for (int v = 0; v < 3; v++)
            if (ms[18 + v] < b32) { b32 = ms[18 + v]; S->fbR32 = kR[v]; S->fbL32 = 0; }
    }
    S->evictions = S->loads = 0;
    S->peak = S->allocated;
    S->t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int i = 0; i < S->nSlots; i++) pool.emplace_back(worker, S, i);
    for (auto &t : pool) t.join();
    if (!S->failure.empty()) throw ServerError(OD_ERR_HIP, "server run failed: " + S->failure);
    for (int k = 0; k < nReq; k++) {
        const Request &r = S->reqs[k];
        const int dst = order[k];
        if (start) start[dst] = r.start;
        if (finish) finish[dst] = r.finish;
        if (iters) iters[dst] = r.iters;
    }
    if (stats) {
        stats[0] = S->peak;
        stats[1] = S->evictions;
        stats[2] = S->loads;
        stats[3] = S->resident;
    }
    SRV_END
}

int od_server_object_state(od_server *S, int objId, double *d, double *v, double *a) {
    SRV_BEGIN
    srv_require(S && objId >= 0 && objId < (int)S->objs.size(), "unknown object");
    Object &o = S->objs[objId];
    if (o.it) {
        check(od_integrator_get_state(o.it, d, v, a), "od_integrator_get_state");
    } else {
        srv_require(o.hasState || o.steps == 0, "object state missing", OD_ERR_STATE);
        const size_t n = (size_t)o.ndof;
        if (o.hasState) {
            if (d) std::memcpy(d, o.d.data(), 8 * n);
            if (v) std::memcpy(v, o.v.data(), 8 * n);
            if (a) std::memcpy(a, o.a.data(), 8 * n);
        } else {
            if (d) std::memset(d, 0, 8 * n);
            if (v) std::memset(v, 0, 8 * n);
            if (a) std::memset(a, 0, 8 * n);
        }
    }
    SRV_END
}

int od_server_object_info(od_server *S, int objId, long long *info) {
    SRV_BEGIN
    srv_require(S && info && objId >= 0 && objId < (int)S->objs.size(), "unknown object");
    const Object &o = S->objs[objId];
    info[0] = o.ndof;
    info[1] = o.mesh ? 1 : 0;
    info[2] = o.bytes;
    info[3] = o.steps;
    info[4] = o.nTets;
    SRV_END
}

}  // extern "C"
