// od_gpu.hpp — the reference's host interfaces, re-declared with the reference's signatures,
// and GPU implementations of them over libodhip.so (od.h).  Host code written against the
// reference's solver, preconditioner, material, integrator and simulator interfaces switches to
// the GPU path by changing the concrete types it constructs.
//
// Interfaces re-declared (paths relative to OpenDeformer/ in the reference tree):
//   ReferenceCounted, Reference<T> ........ core/memory.h:18-76
//   Mesh (accessors), TetMesh .............. core/mesh.h:15-64
//   Constrainer, NodeIndexer ............... core/constrainer.h:16-36, core/nodeIndexer.h:12-29
//   Forcer, StaDampingForcer ............... core/forcer.h:13-26
//   BlockedSymSpMatrix (+ Assembler), SpMDV  core/sparseMatrix.h:186-295, 620-666;
//                                            core/numerMethod.h:134-136
//   SparseSymMatrixIndicesPerElementCache .. core/sparseMatrix.h:650-666
//   LinearSolver<SpMatrix> ................. core/linearSolver.h:12-20
//   Preconditioner ......................... core/preconditioner.h:12-22
//   FullOrderNonlinearMaterialCache,
//   FullOrderNonlinearMaterial<SpMatrix> ... material/fullOrderNonlinear.h:14-44
//   MechMaterial::generateMassMatrix ....... core/mechMaterial.h:29-32
//   Intergrator ............................ core/intergrator.h:13-35 (the damping hook takes a
//                                            StaDampingForcer, as the override at
//                                            nonlinearImplicitBackwardEuler.h:23 does)
//   Simulator .............................. core/simulator.h:12-28, core/simulator.cpp:24-38
//   StaticPlaneShape,
//   MeshPlaneCollisionDampingForcer ........ collision/collisionShapes.h:17-20,
//                                            collision/penaltyDampingForcer.h:25-38
//
// GPU implementations: GpuCGSolver : LinearSolver<BlockedSymSpMatrix>; GpuJacobiPreconditioner,
// GpuIC0Preconditioner, GpuInCholeskyPreconditioner, GpuInLDLTPreconditioner : Preconditioner;
// GpuCorotationalMaterial, GpuNeoHookeanMaterial, GpuPerTetMaterial :
// FullOrderNonlinearMaterial<BlockedSymSpMatrix>; GpuNewmark,
// GpuBackwardEuler : Intergrator.  The reference's class names are aliases of them.
//
// Semantics kept from the reference: host pointers in and out, DOF vectors over free DOFs;
// generateMatrixAndVirtualWorks, generateMassMatrix and SpMDV accumulate into their outputs;
// solveLinearSystem treats `result` as the initial guess and stops when ||r|| <= tol ||b|| or
// after maxIteration iterations; the preconditioner is rebuilt from the matrix at every solve;
// errors are reported, not thrown (the reference's Severe prints and continues,
// core/oder.h:117-119): here they are printed with the library's error text and recorded in
// lastStatus().  Differences: the device runs one HIP stream per process-wide Device; matrix
// values live on the device and are mirrored to the host on demand (hostValues /
// markHostDirty); constraints must cover whole nodes; materials may differ per tet.
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <utility>
#include <vector>

#include "od.h"

namespace odgpu {

using Scalar = double;

inline int Severe(int status, const char *where) {
    if (status != OD_OK) std::fprintf(stderr, "odgpu: %s: %s\n", where, od_last_error());
    return status;
}

// ---- reference counting (core/memory.h:18-76) ------------------------------------------
class ReferenceCounted {
protected:
    ReferenceCounted() : nReference(0) {}
    std::atomic<int> nReference;
    template <class T> friend class Reference;
};

template <class T> class Reference {
public:
    Reference(T *ptr = nullptr) noexcept : p(ptr) { if (p) p->nReference.fetch_add(1, std::memory_order_relaxed); }
    Reference(const Reference &r) noexcept : p(r.p) { if (p) p->nReference.fetch_add(1, std::memory_order_relaxed); }
    Reference(Reference &&r) noexcept : p(r.p) { r.p = nullptr; }
    template <class U> Reference(const Reference<U> &r) noexcept : p(r.get()) {
        if (p) p->nReference.fetch_add(1, std::memory_order_relaxed);
    }
    Reference &operator=(T *ptr) noexcept {
        if (ptr) ptr->nReference.fetch_add(1, std::memory_order_relaxed);
        release();
        p = ptr;
        return *this;
    }
    Reference &operator=(const Reference &r) noexcept { return *this = r.p; }
    Reference &operator=(Reference &&r) noexcept {
        std::swap(p, r.p);
        return *this;
    }
    ~Reference() { release(); }
    T &operator*() const noexcept { return *p; }
    T *operator->() noexcept { return p; }
    const T *operator->() const noexcept { return p; }
    operator bool() const noexcept { return p != nullptr; }
    T *get() const noexcept { return p; }

private:
    void release() noexcept {
        if (p && p->nReference.fetch_sub(1, std::memory_order_acq_rel) == 1) delete p;
    }
    T *p;
};

struct Vector3 {
    Scalar x = 0, y = 0, z = 0;
    Vector3() = default;
    Vector3(Scalar a, Scalar b, Scalar c) : x(a), y(b), z(c) {}
    Scalar operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Scalar &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    Vector3 operator+(const Vector3 &o) const { return Vector3(x + o.x, y + o.y, z + o.z); }
};
struct Vector3f {
    float x = 0, y = 0, z = 0;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

// ---- the device (one context, one HIP stream) ---------------------------------------------
// The process-wide device the classes below run on (device 0 unless selectDevice is called
// before first use).
class Device {
public:
    static Device &instance() {
        static Device d(selected());
        return d;
    }
    static void selectDevice(int device) { selected() = device; }
    od_ctx *get() const { return ctx_; }
    int lastStatus() const { return status_; }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    ~Device() { if (ctx_) od_ctx_destroy(ctx_); }

private:
    explicit Device(int device) { status_ = Severe(od_ctx_create(device, &ctx_), "od_ctx_create"); }
    static int &selected() {
        static int d = 0;
        return d;
    }
    od_ctx *ctx_ = nullptr;
    int status_ = OD_OK;
};

// Device buffer of T, grown on demand.
template <typename T> class DeviceBuffer {
public:
    DeviceBuffer() = default;
    ~DeviceBuffer() { if (p_) od_dev_free(Device::instance().get(), p_); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    T *ensure(long long n) {
        if (n <= n_) return p_;
        od_ctx *c = Device::instance().get();
        if (p_) od_dev_free(c, p_);
        p_ = nullptr;
        n_ = 0;
        void *q = nullptr;
        if (Severe(od_dev_alloc(c, n * (long long)sizeof(T), &q), "od_dev_alloc") == OD_OK) {
            p_ = (T *)q;
            n_ = n;
        }
        return p_;
    }
    int upload(const T *host, long long n) {
        if (!ensure(std::max<long long>(n, 1))) return OD_ERR_HIP;
        return Severe(od_copy_h2d(Device::instance().get(), p_, host, n * (long long)sizeof(T)), "od_copy_h2d");
    }
    int download(T *host, long long n) const {
        return Severe(od_copy_d2h(Device::instance().get(), host, p_, n * (long long)sizeof(T)), "od_copy_d2h");
    }
    T *get() const { return p_; }

private:
    T *p_ = nullptr;
    long long n_ = 0;
};

// ---- mesh (core/mesh.h:15-64) ------------------------------------------------------------
class Mesh : public ReferenceCounted {
public:
    Mesh(int numN, int numE, int numS, int numNPE, int numVPS = 3)
        : numNodes(numN), numElements(numE), numSurfaces(numS), numNodesPerElement(numNPE), numVertPerSur(numVPS),
          vertices((size_t)numN), elements((size_t)numE * numNPE), surfaces((size_t)numS * numVPS),
          displacements((size_t)numN) {}
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;
    virtual ~Mesh() = default;
    int getNodeCount() const { return numNodes; }
    int getNodePerElementCount() const { return numNodesPerElement; }
    int getElementCount() const { return numElements; }
    int getVertPerFacetCount() const { return numVertPerSur; }
    int getFacetCount() const { return numSurfaces; }
    const Vector3 &getVertex(int vertIndex) const { return vertices[vertIndex]; }
    Vector3 &getVertexDisplacement(int vertIndex) { return displacements[vertIndex]; }
    const Vector3 &getVertexDisplacementConst(int vertIndex) const { return displacements[vertIndex]; }
    void setVertex(int vertIndex, const Vector3 &v) { vertices[vertIndex] = v; }
    int getElementNodeIndex(int elementIndex, int nodeIndex) { return elements[elementIndex * numNodesPerElement + nodeIndex]; }
    const int *getElementNodeReference(int elementIndex) const { return &elements[(size_t)elementIndex * numNodesPerElement]; }
    const int *getFacetVertReference(int facetIndex) const { return &surfaces[(size_t)facetIndex * numVertPerSur]; }
    void setElement(int elementIndex, const int *nodeIndices) {
        for (int i = 0; i < numNodesPerElement; i++) elements[(size_t)elementIndex * numNodesPerElement + i] = nodeIndices[i];
    }
    void setFacet(int facetIndex, const int *facetIndices) {
        for (int i = 0; i < numVertPerSur; i++) surfaces[(size_t)facetIndex * numVertPerSur + i] = facetIndices[i];
    }

private:
    int numNodes, numElements, numSurfaces;
    const int numNodesPerElement, numVertPerSur;
    std::vector<Vector3> vertices;
    std::vector<int> elements, surfaces;
    std::vector<Vector3> displacements;
};

// A tetrahedral mesh from arrays (the reference's TetMesh is read from files): 3 coordinates
// per node, 4 nodes per tet, optionally 3 vertices per boundary facet.
class TetMesh : public Mesh {
public:
    TetMesh(int nNodes, const Scalar *X, int nTets, const int *tets, int nFacets = 0, const int *facets = nullptr)
        : Mesh(nNodes, nTets, nFacets, 4, 3) {
        for (int n = 0; n < nNodes; n++) setVertex(n, Vector3(X[3 * n], X[3 * n + 1], X[3 * n + 2]));
        for (int e = 0; e < nTets; e++) setElement(e, tets + 4 * (size_t)e);
        for (int f = 0; f < nFacets; f++) setFacet(f, facets + 3 * (size_t)f);
    }
};

// ---- constraints and DOF indexing (core/constrainer.h:16-36, core/nodeIndexer.h:12-29) ---------
enum ConstrainType { Constrain_X = 1, Constrain_Y = 2, Constrain_Z = 4, Constrain_XY = 3, Constrain_YZ = 6,
                     Constrain_XZ = 5, Constrain_XYZ = 7 };

class Constrainer {
public:
    Constrainer(const Reference<Mesh> &m) : mesh(m) {}
    void setConstrains(int nodeCounts, int *nodeIndices, ConstrainType *constrainTypes = nullptr) {
        for (int i = 0; i < nodeCounts; i++) setConstrain(nodeIndices[i], constrainTypes ? constrainTypes[i] : Constrain_XYZ);
    }
    // every vertex of a facet with |(v - o) . n| < 1e-4 (core/constrainer.cpp:57-72)
    void setConstrains(const Vector3 &origin, const Vector3 &normal, ConstrainType constrainType = Constrain_XYZ) {
        const int vpf = mesh->getVertPerFacetCount();
        for (int f = 0; f < mesh->getFacetCount(); f++) {
            const int *v = mesh->getFacetVertReference(f);
            bool on = true;
            for (int k = 0; k < vpf && on; k++) {
                const Vector3 &p = mesh->getVertex(v[k]);
                const Scalar d = (p.x - origin.x) * normal.x + (p.y - origin.y) * normal.y + (p.z - origin.z) * normal.z;
                on = std::fabs(d) < 1e-4;
            }
            if (on)
                for (int k = 0; k < vpf; k++) setConstrain(v[k], constrainType);
        }
    }
    void outputConstrains(std::vector<int> &out) const { out.assign(constrains.begin(), constrains.end()); }
    const Reference<Mesh> &getMesh() const { return mesh; }

private:
    void setConstrain(int nodeIndex, ConstrainType t) {
        for (int a = 0; a < 3; a++)
            if (t & (1 << a)) constrains.insert(3 * nodeIndex + a);
    }
    std::set<int> constrains;
    Reference<Mesh> mesh;
};

class BlockedSymSpMatrix;

class NodeIndexer : public ReferenceCounted {
public:
    typedef std::vector<int>::const_iterator ConstrainIter;
    NodeIndexer(const Constrainer &constrainer) : mesh_(constrainer.getMesh()) { constrainer.outputConstrains(constrainIndices); }
    ~NodeIndexer() { if (dev_) od_mesh_destroy(dev_); }
    // core/nodeIndexer.cpp:12-25: -1 for a constrained DOF, else its compacted index
    int getGlobalIndex(int nodeIndex, int axis) const {
        const int dof = 3 * nodeIndex + axis;
        auto p = std::lower_bound(constrainIndices.begin(), constrainIndices.end(), dof);
        if (p != constrainIndices.end() && *p == dof) return -1;
        return dof - (int)(p - constrainIndices.begin());
    }
    void getElementNodesGlobalIndices(const Reference<Mesh> &mesh, int elementIndex, int nodeCounts, int *nodeIndices) const {
        const int *e = mesh->getElementNodeReference(elementIndex);
        for (int i = 0; i < nodeCounts; i++)
            for (int a = 0; a < 3; a++) nodeIndices[3 * i + a] = getGlobalIndex(e[i], a);
    }
    int getMatrixOrder(const Reference<Mesh> &mesh) const { return 3 * mesh->getNodeCount() - (int)constrainIndices.size(); }
    ConstrainIter getConstrainIterBegin() const { return constrainIndices.begin(); }
    ConstrainIter getConstrainIterEnd() const { return constrainIndices.end(); }

    // The device copy of (mesh, constraints): pattern, maps, level schedule (od_mesh_create),
    // built on first use and shared by every matrix, material and integrator of this indexer.
    od_mesh *deviceMesh() const {
        if (!dev_) {
            const int nn = mesh_->getNodeCount(), ne = mesh_->getElementCount();
            std::vector<Scalar> X(3 * (size_t)nn);
            for (int n = 0; n < nn; n++)
                for (int a = 0; a < 3; a++) X[3 * (size_t)n + a] = mesh_->getVertex(n)[a];
            std::vector<int> tets(4 * (size_t)ne);
            for (int e = 0; e < ne; e++) std::memcpy(&tets[4 * (size_t)e], mesh_->getElementNodeReference(e), 4 * sizeof(int));
            status_ = Severe(od_mesh_create(Device::instance().get(), nn, X.data(), ne, tets.data(),
                                            (int)constrainIndices.size(), constrainIndices.data(), 1, nullptr, nullptr,
                                            &dev_),
                             "od_mesh_create");
        }
        return dev_;
    }
    // free-DOF vector of a per-node quantity, and back (constrained DOFs read as 0)
    void gatherFree(const Reference<Mesh> &mesh, const Vector3 *perNode, Scalar *freeDofs) const {
        for (int n = 0; n < mesh->getNodeCount(); n++)
            for (int a = 0; a < 3; a++) {
                const int g = getGlobalIndex(n, a);
                if (g >= 0) freeDofs[g] = perNode[n][a];
            }
    }
    void scatterFree(const Reference<Mesh> &mesh, const Scalar *freeDofs, Vector3 *perNode) const {
        for (int n = 0; n < mesh->getNodeCount(); n++)
            for (int a = 0; a < 3; a++) {
                const int g = getGlobalIndex(n, a);
                perNode[n][a] = g >= 0 ? freeDofs[g] : Scalar(0);
            }
    }
    std::vector<Scalar> freeDisplacements(const Reference<Mesh> &mesh) const {
        std::vector<Scalar> u((size_t)getMatrixOrder(mesh), 0.0);
        std::vector<Vector3> d((size_t)mesh->getNodeCount());
        for (int n = 0; n < mesh->getNodeCount(); n++) d[n] = mesh->getVertexDisplacementConst(n);
        gatherFree(mesh, d.data(), u.data());
        return u;
    }
    int lastStatus() const { return status_; }

private:
    Reference<Mesh> mesh_;
    std::vector<int> constrainIndices;
    mutable od_mesh *dev_ = nullptr;
    mutable int status_ = OD_OK;
};

// ---- loads (core/forcer.h:13-26, core/forcer.cpp:8-96) ---------------------------------------
class Forcer {
public:
    // V/4 * bodyForce on each node of every tet (element/tetelement.cpp:68-77)
    void addBodyForce(const Reference<Mesh> &mesh, Scalar bodyForce[3], const Reference<NodeIndexer> &indexer) {
        int gi[12];
        for (int e = 0; e < mesh->getElementCount(); e++) {
            const int *t = mesh->getElementNodeReference(e);
            const Vector3 &a = mesh->getVertex(t[0]), &b = mesh->getVertex(t[1]), &c = mesh->getVertex(t[2]),
                          &d = mesh->getVertex(t[3]);
            // getTetVolume (element/tetelement.cpp:6-12): |ab . (ac x ad)| / 6
            const Scalar ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z, vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
            const Scalar wx = d.x - a.x, wy = d.y - a.y, wz = d.z - a.z;
            const Scalar cx = vy * wz - vz * wy, cy = vz * wx - vx * wz, cz = vx * wy - vy * wx;
            const Scalar volume = std::fabs(ux * cx + uy * cy + uz * cz) / Scalar(6.0);
            const Scalar factor = volume / Scalar(4.0);
            indexer->getElementNodesGlobalIndices(mesh, e, 4, gi);
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 3; k++) {
                    const Scalar r = factor * bodyForce[k];
                    if (r != 0.0 && gi[3 * j + k] >= 0) virtualWorks[gi[3 * j + k]] += r;
                }
        }
    }
    void addSurfaceForceByNode(const Reference<Mesh> &mesh, Scalar nodeForce[3], int nodeCounts, int *nodeIndex,
                               const Reference<NodeIndexer> &indexer) {
        (void)mesh;
        for (int i = 0; i < nodeCounts; i++)
            for (int j = 0; j < 3; j++) {
                const int g = nodeForce[j] != 0.0 ? indexer->getGlobalIndex(nodeIndex[i], j) : -1;
                if (g >= 0) virtualWorks[g] += nodeForce[j];
            }
    }
    // basises == NULL: the full-order virtual works (forcer.cpp:80-96)
    void getVirtualWorks(int dofs, int totalDofs, const Scalar *basises, Scalar *vws) const {
        std::memset(vws, 0, sizeof(Scalar) * (size_t)dofs);
        if (basises) {
            const Scalar *basis = basises;
            for (int i = 0; i < dofs; i++, basis += totalDofs)
                for (const auto &e : virtualWorks) vws[i] += e.second * basis[e.first];
        } else {
            for (const auto &e : virtualWorks) vws[e.first] = e.second;
        }
    }

private:
    std::map<int, Scalar> virtualWorks;
};

// ---- the blocked symmetric matrix (core/sparseMatrix.h:186-295) -----------------------------
// Per-element 78-entry value indices (core/sparseMatrix.h:650-666), as getMatrixIndicesPerElement
// returns them; the device keeps its own block-level form of the same map.
class SparseSymMatrixIndicesPerElementCache {
public:
    SparseSymMatrixIndicesPerElementCache() = default;
    SparseSymMatrixIndicesPerElementCache(int elementCount, int nodePerElementCount)
        : count_(nodePerElementCount * 3 * (nodePerElementCount * 3 + 1) / 2),
          indices_((size_t)elementCount * count_, -1) {}
    SparseSymMatrixIndicesPerElementCache(SparseSymMatrixIndicesPerElementCache &&) = default;
    SparseSymMatrixIndicesPerElementCache &operator=(SparseSymMatrixIndicesPerElementCache &&) = default;
    const int *getElementMatIndices(int elementIndex) const { return indices_.data() + (size_t)elementIndex * count_; }
    int *getElementMatIndices(int elementIndex) { return indices_.data() + (size_t)elementIndex * count_; }

private:
    int count_ = 0;
    std::vector<int> indices_;
};

// Values live on the device in the reference's canonical layout; hostValues() mirrors them and
// markHostDirty() pushes host edits back before the next device use.
class BlockedSymSpMatrix {
public:
    // BlockedSymSparseMatrixAssembler: getMatrixStructure binds it to (mesh, indexer); the device
    // derives the pattern itself (od_mesh_create), identical to the reference's.
    struct Assembler {
        explicit Assembler(int numColumns) : numColumn(numColumns) {}
        void addEntry(int, int, Scalar) {}
        int numColumn;
        Reference<NodeIndexer> indexer;
    };
    explicit BlockedSymSpMatrix(const Assembler &assembler) : indexer_(assembler.indexer) {
        if (!indexer_) { status_ = OD_ERR_STATE; std::fprintf(stderr, "odgpu: matrix without a structure\n"); return; }
        status_ = Severe(od_matrix_create(indexer_->deviceMesh(), &m_), "od_matrix_create");
        own_ = true;
        describe();
    }
    // a view of a matrix owned elsewhere (the integrator's damping matrix)
    BlockedSymSpMatrix(const Reference<NodeIndexer> &indexer, od_matrix *m) : indexer_(indexer), m_(m) { describe(); }
    ~BlockedSymSpMatrix() { if (own_ && m_) od_matrix_destroy(m_); }
    BlockedSymSpMatrix(const BlockedSymSpMatrix &) = delete;
    BlockedSymSpMatrix &operator=(const BlockedSymSpMatrix &) = delete;

    int getNumColumns() const { return numColumns_; }
    int getNumEntries() const { return (int)nValues_; }
    void setZeros() { host_.clear(); hostDirty_ = false; status_ = Severe(od_matrix_zero(m_), "od_matrix_zero"); }
    void Scale(Scalar alpha) {
        deviceSync();
        status_ = Severe(od_matrix_axpby(m_, alpha, 0.0, nullptr), "od_matrix_axpby");
        host_.clear();
    }
    void Add(Scalar alpha, const BlockedSymSpMatrix &mat) {
        deviceSync();
        mat.deviceSync();
        status_ = Severe(od_matrix_axpby(m_, 1.0, alpha, mat.m_), "od_matrix_axpby");
        host_.clear();
    }
    // BlockedSymSparseMatrix::getIndices as getMatrixIndicesPerElement uses it: the per-tet map
    SparseSymMatrixIndicesPerElementCache elementIndices(int elementCount) const {
        SparseSymMatrixIndicesPerElementCache c(elementCount, 4);
        if (elementCount > 0)
            Severe(od_pattern_export(indexer_->deviceMesh(), nullptr, nullptr, nullptr, c.getElementMatIndices(0)),
                   "od_pattern_export");
        return c;
    }
    std::vector<Scalar> &hostValues() {
        if (host_.empty()) {
            host_.assign((size_t)nValues_, 0.0);
            status_ = Severe(od_matrix_download(m_, host_.data()), "od_matrix_download");
        }
        return host_;
    }
    void markHostDirty() { hostDirty_ = true; }
    void deviceSync() const {
        if (hostDirty_ && !host_.empty()) {
            Severe(od_matrix_upload(m_, host_.data()), "od_matrix_upload");
            hostDirty_ = false;
        }
    }
    void deviceWritten() { host_.clear(); hostDirty_ = false; }
    od_matrix *get() const { return m_; }
    const Reference<NodeIndexer> &indexer() const { return indexer_; }
    int lastStatus() const { return status_; }

private:
    void describe() {
        long long info[16] = {0};
        if (indexer_ && indexer_->deviceMesh()) od_mesh_info(indexer_->deviceMesh(), info);
        numColumns_ = (int)info[0];
        nValues_ = info[3];
    }
    Reference<NodeIndexer> indexer_;
    od_matrix *m_ = nullptr;
    bool own_ = false;
    int numColumns_ = 0;
    long long nValues_ = 0;
    mutable std::vector<Scalar> host_;
    mutable bool hostDirty_ = false;
    int status_ = OD_OK;
};

// SpMDV (core/numerMethod.h:134-136): dest += mat * src, host vectors.
inline int SpMDV(const BlockedSymSpMatrix &mat, const Scalar *src, Scalar *dest) {
    const int n = mat.getNumColumns();
    DeviceBuffer<Scalar> x, y;
    mat.deviceSync();
    int st = x.upload(src, n);
    if (st == OD_OK) st = y.upload(dest, n);
    if (st == OD_OK) st = Severe(od_spmv(mat.get(), x.get(), y.get(), 1), "od_spmv");
    if (st == OD_OK) st = y.download(dest, n);
    return st;
}

// ---- linear solvers and preconditioners (core/linearSolver.h:12-20, core/preconditioner.h:12-22) --
template <class SpMatrix> class LinearSolver : public ReferenceCounted {
public:
    LinearSolver(const SpMatrix *m = nullptr) : mat(m) {}
    virtual void resetLinearSystem(const SpMatrix *m) { mat = m; }
    virtual void solveLinearSystem(const Scalar *rhs, Scalar *result) = 0;
    virtual ~LinearSolver() = default;

protected:
    const SpMatrix *mat;
};

class Preconditioner : public ReferenceCounted {
public:
    // the identity (the reference's default Preconditioner)
    virtual void solvePreconditionerSystem(int width, const Scalar *rhs, Scalar *result) const {
        std::memcpy(result, rhs, (size_t)width * sizeof(Scalar));
    }
    virtual void resetPreconditionerSystem(const BlockedSymSpMatrix &mat) { bound_ = &mat; }
    // builds the factor (or D^-1) of mat on the device; solvePreconditionerSystem applies it
    virtual void Preprocess(const BlockedSymSpMatrix &mat) {
        bound_ = &mat;
        if (deviceKind() == OD_PRE_NONE || !ensureSolver()) return;
        mat.deviceSync();
        Severe(od_precond_preprocess(s_, mat.get()), "od_precond_preprocess");
    }
    virtual ~Preconditioner() { if (s_) od_solver_destroy(s_); }
    // the library's kind and parameters of this preconditioner (OD_PRE_NONE for the identity)
    virtual int deviceKind() const { return OD_PRE_NONE; }
    virtual Scalar deviceParam1() const { return 0.0; }
    virtual Scalar deviceParam2() const { return 0.0; }

protected:
    bool ensureSolver() const {
        if (s_) return true;
        if (!bound_) { std::fprintf(stderr, "odgpu: preconditioner without a matrix (Preprocess first)\n"); return false; }
        if (Severe(od_solver_create(bound_->indexer()->deviceMesh(), deviceKind(), deviceParam1(), OD_FP64, &s_),
                   "od_solver_create"))
            return false;
        return Severe(od_solver_set_precond_params(s_, deviceParam1(), deviceParam2()),
                      "od_solver_set_precond_params") == OD_OK;
    }
    // the device apply of the subclasses, with the factor of the last Preprocess
    void deviceApply(int width, const Scalar *rhs, Scalar *result) const {
        if (!ensureSolver()) return;
        DeviceBuffer<Scalar> r, z;
        if (r.upload(rhs, width) || !z.ensure(width)) return;
        if (Severe(od_precond_solve(s_, r.get(), z.get()), "od_precond_solve")) return;
        z.download(result, width);
    }
    const BlockedSymSpMatrix *bound_ = nullptr;
    mutable od_solver *s_ = nullptr;
};

// JacobiPreconditioner (solver/jacobiPreconditioner.h:10-22)
class GpuJacobiPreconditioner : public Preconditioner {
public:
    void solvePreconditionerSystem(int width, const Scalar *rhs, Scalar *result) const override {
        deviceApply(width, rhs, result);
    }
    int deviceKind() const override { return OD_PRE_JACOBI; }
};

// IC(0) on the block pattern, the device-factored incomplete Cholesky; epsilon is the pivot
// guard d <= 0 -> epsilon^2 d^2 (solver/inCholeskyPreconditioner.cpp:52).
class GpuIC0Preconditioner : public Preconditioner {
public:
    explicit GpuIC0Preconditioner(Scalar epsilon = 1e-3) : eps_(epsilon) {}
    void solvePreconditionerSystem(int width, const Scalar *rhs, Scalar *result) const override {
        deviceApply(width, rhs, result);
    }
    int deviceKind() const override { return OD_PRE_IC0; }
    Scalar deviceParam1() const override { return eps_; }

private:
    Scalar eps_;
};

// InCholeskyPreconditioner(epsilon, relaxedScale) (solver/inCholeskyPreconditioner.h:14)
class GpuInCholeskyPreconditioner : public Preconditioner {
public:
    GpuInCholeskyPreconditioner(Scalar epsilon, Scalar relaxedScale) : eps_(epsilon), relax_(relaxedScale) {}
    void solvePreconditionerSystem(int width, const Scalar *rhs, Scalar *result) const override {
        deviceApply(width, rhs, result);
    }
    int deviceKind() const override { return OD_PRE_ICT; }
    Scalar deviceParam1() const override { return eps_; }
    Scalar deviceParam2() const override { return relax_; }

private:
    Scalar eps_, relax_;
};

// InLDLTPreconditioner(sainvEpsilon, ldltEpsilon) (solver/inLDLTPreconditioner.h:14-15)
class GpuInLDLTPreconditioner : public Preconditioner {
public:
    GpuInLDLTPreconditioner(Scalar sainvEpsilon, Scalar ldltEpsilon) : se_(sainvEpsilon), le_(ldltEpsilon) {}
    void solvePreconditionerSystem(int width, const Scalar *rhs, Scalar *result) const override {
        deviceApply(width, rhs, result);
    }
    int deviceKind() const override { return OD_PRE_LDLT; }
    Scalar deviceParam1() const override { return se_; }
    Scalar deviceParam2() const override { return le_; }

private:
    Scalar se_, le_;
};

// CGSolver(tol, maxIter, preconditioner, matrix) (solver/cgSolver.h:14-30).  `precision`
// (no reference counterpart) selects the FP32-inner form.
class GpuCGSolver : public LinearSolver<BlockedSymSpMatrix> {
public:
    GpuCGSolver(Scalar tol, int maxIter, const Reference<Preconditioner> &precondition = new Preconditioner(),
                const BlockedSymSpMatrix *m = nullptr, int precision = OD_FP64)
        : LinearSolver(m), tolerant(tol), maxIteration(maxIter), preconditioner(precondition), precision_(precision) {}
    GpuCGSolver(const GpuCGSolver &) = delete;
    GpuCGSolver &operator=(const GpuCGSolver &) = delete;
    ~GpuCGSolver() override { if (s_) od_solver_destroy(s_); }
    void resetLinearSystem(const BlockedSymSpMatrix *m) override {
        if (s_ && m && mat && m->indexer().get() != mat->indexer().get()) { od_solver_destroy(s_); s_ = nullptr; }
        mat = m;
    }
    // result: initial guess in, solution out (solver/cgSolver.cpp:13-57)
    void solveLinearSystem(const Scalar *rhs, Scalar *result) override {
        if (!mat) { status_ = OD_ERR_STATE; std::fprintf(stderr, "odgpu: CGSolver without a matrix\n"); return; }
        const int n = mat->getNumColumns();
        if (!s_) {
            const Preconditioner &p = *preconditioner;
            status_ = Severe(od_solver_create(mat->indexer()->deviceMesh(), p.deviceKind(), p.deviceParam1(), precision_,
                                              &s_),
                             "od_solver_create");
            if (status_) return;
            status_ = Severe(od_solver_set_precond_params(s_, p.deviceParam1(), p.deviceParam2()),
                             "od_solver_set_precond_params");
            if (status_) return;
        }
        mat->deviceSync();
        if ((status_ = b_.upload(rhs, n)) || (status_ = x_.upload(result, n))) return;
        status_ = Severe(od_pcg_solve(s_, mat->get(), b_.get(), x_.get(), tolerant, maxIteration, &iters_, &relres_),
                         "od_pcg_solve");
        if (status_ == OD_OK) status_ = x_.download(result, n);
    }
    Scalar tolerance() const { return tolerant; }
    int maxIterations() const { return maxIteration; }
    const Reference<Preconditioner> &getPreconditioner() const { return preconditioner; }
    int precision() const { return precision_; }
    int lastIterations() const { return iters_; }
    Scalar lastRelativeResidual() const { return relres_; }
    int lastStatus() const { return status_; }

private:
    Scalar tolerant;
    int maxIteration;
    Reference<Preconditioner> preconditioner;
    int precision_;
    od_solver *s_ = nullptr;
    DeviceBuffer<Scalar> b_, x_;
    int iters_ = 0;
    Scalar relres_ = 0;
    int status_ = OD_OK;
};

// ---- materials (material/fullOrderNonlinear.h:14-44, core/mechMaterial.h:29-32) -------------
// The reference's per-material precomputes (K0 and Dm^-1 per tet, material/corotationalHyperelastic.cpp:
// 20-41) are recomputed on the device inside the element kernel, so the cache is empty.
struct FullOrderNonlinearMaterialCache {
    FullOrderNonlinearMaterialCache() = default;
    FullOrderNonlinearMaterialCache(FullOrderNonlinearMaterialCache &&) = default;
    FullOrderNonlinearMaterialCache &operator=(FullOrderNonlinearMaterialCache &&) = default;
    FullOrderNonlinearMaterialCache(const FullOrderNonlinearMaterialCache &) = delete;
    FullOrderNonlinearMaterialCache &operator=(const FullOrderNonlinearMaterialCache &) = delete;
};

template <class SpMatrix> class FullOrderNonlinearMaterial : public ReferenceCounted {
public:
    explicit FullOrderNonlinearMaterial(Scalar rho) : density(rho) {}
    virtual ~FullOrderNonlinearMaterial() = default;
    virtual FullOrderNonlinearMaterialCache getPrecomputes(const Reference<Mesh> &mesh) const {
        (void)mesh;
        return FullOrderNonlinearMaterialCache();
    }
    // K += K(u), vws += f_int(u) at the mesh's current displacements (accumulating)
    virtual void generateMatrixAndVirtualWorks(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer,
                                               const FullOrderNonlinearMaterialCache &cache,
                                               const SparseSymMatrixIndicesPerElementCache &matrixIndices,
                                               SpMatrix &matrix, Scalar *vws) const = 0;
    void getMatrixStructure(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer,
                            typename SpMatrix::Assembler &assembler) const {
        (void)mesh;
        assembler.indexer = indexer;
    }
    SparseSymMatrixIndicesPerElementCache getMatrixIndicesPerElement(const Reference<Mesh> &mesh,
                                                                     const Reference<NodeIndexer> &indexer,
                                                                     const SpMatrix &mat) const {
        (void)indexer;
        return mat.elementIndices(mesh->getElementCount());
    }
    // MechMaterial::generateMassMatrix (core/mechMaterial.cpp:69-98): matrix += M
    void generateMassMatrix(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer,
                            const SparseSymMatrixIndicesPerElementCache &matrixIndices, SpMatrix &matrix) const {
        (void)matrixIndices;
        if (bindTo(mesh, indexer)) return;
        matrix.deviceSync();
        Severe(od_mass_matrix(indexer->deviceMesh(), matrix.get()), "od_mass_matrix");
        matrix.deviceWritten();
    }
    Scalar getDensity() const { return density; }
    // the per-tet device tables (kind, E, nu, rho) of this material on a mesh
    virtual void tables(int nTets, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                        std::vector<Scalar> &rho) const = 0;
    virtual Scalar corotationalThreshold() const { return 0.0; }
    virtual Scalar inversionThreshold() const { return 0.1; }
    int bindTo(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer) const {
        std::vector<int> kind;
        std::vector<Scalar> E, nu, rho;
        tables(mesh->getElementCount(), kind, E, nu, rho);
        od_mesh *dm = indexer->deviceMesh();
        if (!dm) return OD_ERR_STATE;
        return Severe(od_mesh_set_material(dm, kind.data(), E.data(), nu.data(), rho.data(), corotationalThreshold(),
                                           inversionThreshold()),
                      "od_mesh_set_material");
    }

protected:
    Scalar density;
    int assemble(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer, SpMatrix &matrix, Scalar *vws) const {
        int st = bindTo(mesh, indexer);
        if (st) return st;
        const int n = indexer->getMatrixOrder(mesh);
        std::vector<Scalar> u = indexer->freeDisplacements(mesh), f((size_t)n, 0.0);
        DeviceBuffer<Scalar> du, df;
        if ((st = du.upload(u.data(), n)) || (st = df.upload(f.data(), n))) return st;
        matrix.deviceSync();
        st = Severe(od_assemble(indexer->deviceMesh(), du.get(), matrix.get(), df.get()), "od_assemble");
        matrix.deviceWritten();
        if (st == OD_OK) st = df.download(f.data(), n);
        if (st == OD_OK)
            for (int i = 0; i < n; i++) vws[i] += f[(size_t)i];
        return st;
    }
};

template <int KIND> class GpuUniformMaterial : public FullOrderNonlinearMaterial<BlockedSymSpMatrix> {
public:
    GpuUniformMaterial(Scalar rho, Scalar youngs, Scalar possion, Scalar threshold)
        : FullOrderNonlinearMaterial(rho), E_(youngs), nu_(possion), thr_(threshold) {}
    void generateMatrixAndVirtualWorks(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer,
                                       const FullOrderNonlinearMaterialCache &, const SparseSymMatrixIndicesPerElementCache &,
                                       BlockedSymSpMatrix &matrix, Scalar *vws) const override {
        assemble(mesh, indexer, matrix, vws);
    }
    void tables(int n, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                std::vector<Scalar> &rho) const override {
        kind.assign((size_t)n, KIND);
        E.assign((size_t)n, E_);
        nu.assign((size_t)n, nu_);
        rho.assign((size_t)n, density);
    }
    Scalar corotationalThreshold() const override { return KIND == OD_MAT_NEOHOOKEAN ? 0.0 : thr_; }
    Scalar inversionThreshold() const override { return KIND == OD_MAT_NEOHOOKEAN ? thr_ : 0.1; }

private:
    Scalar E_, nu_, thr_;
};
// CorotationalHyperelasticMaterial(rho, youngs, possion, compressingTreshold)
// (material/corotationalHyperelastic.h:13)
using GpuCorotationalMaterial = GpuUniformMaterial<OD_MAT_COROTATIONAL>;
// InvertibleNeoHookeanMaterial(rho, youngs, possion, inversionTrashold) (material/invertibleNeoHookean.h:13)
using GpuNeoHookeanMaterial = GpuUniformMaterial<OD_MAT_NEOHOOKEAN>;

// Per-tet material tables (no reference counterpart: the reference binds one material per mesh).
class GpuPerTetMaterial : public FullOrderNonlinearMaterial<BlockedSymSpMatrix> {
public:
    GpuPerTetMaterial(std::vector<int> kind, std::vector<Scalar> E, std::vector<Scalar> nu, std::vector<Scalar> rho,
                      Scalar corotThreshold = 0.0, Scalar neoThreshold = 0.1)
        : FullOrderNonlinearMaterial(rho.empty() ? 0.0 : rho[0]), kind_(std::move(kind)), E_(std::move(E)),
          nu_(std::move(nu)), rho_(std::move(rho)), ct_(corotThreshold), nt_(neoThreshold) {}
    void generateMatrixAndVirtualWorks(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer,
                                       const FullOrderNonlinearMaterialCache &, const SparseSymMatrixIndicesPerElementCache &,
                                       BlockedSymSpMatrix &matrix, Scalar *vws) const override {
        assemble(mesh, indexer, matrix, vws);
    }
    void tables(int, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                std::vector<Scalar> &rho) const override {
        kind = kind_; E = E_; nu = nu_; rho = rho_;
    }
    Scalar corotationalThreshold() const override { return ct_; }
    Scalar inversionThreshold() const override { return nt_; }

private:
    std::vector<int> kind_;
    std::vector<Scalar> E_, nu_, rho_;
    Scalar ct_, nt_;
};

// ---- collision damping (collision/collisionShapes.h:17-20, collision/penaltyDampingForcer.h:25-38) --
class StaDampingForcer {
public:
    virtual void addDampingMatrix(const SparseSymMatrixIndicesPerElementCache &matrixIndices,
                                  BlockedSymSpMatrix &mat) const = 0;
    virtual ~StaDampingForcer() = default;
};

struct StaticPlaneShape {
    Vector3 planeNormal;   // below the plane: n.x + c < 0
    Scalar planeConstant = 0;
};

// The reference's forcer holds one (plane, tet) pair; addCollisionPair (no reference
// counterpart) queues more, added in order by one addDampingMatrix call.  Positions are the
// rest positions plus the mesh's displacements; velocities are per node, as the caller fills
// them from Intergrator::getMeshVerticesVelocities.
class MeshPlaneCollisionDampingForcer : public StaDampingForcer {
public:
    MeshPlaneCollisionDampingForcer(Scalar dampingCoefficient, const Reference<Mesh> &mesh,
                                    const Reference<NodeIndexer> &indexer, const Vector3 *velocities)
        : coeff_(dampingCoefficient), mesh_(mesh), indexer_(indexer), velocities_(velocities) {}
    void setCollisionPair(const StaticPlaneShape &plane, int elementIndex) {
        planes_.clear();
        tets_.clear();
        addCollisionPair(plane, elementIndex);
    }
    void addCollisionPair(const StaticPlaneShape &plane, int elementIndex) {
        planes_.insert(planes_.end(), {plane.planeNormal.x, plane.planeNormal.y, plane.planeNormal.z, plane.planeConstant});
        tets_.push_back(elementIndex);
    }
    void addDampingMatrix(const SparseSymMatrixIndicesPerElementCache &, BlockedSymSpMatrix &mat) const override {
        const int n = indexer_->getMatrixOrder(mesh_);
        std::vector<Scalar> u = indexer_->freeDisplacements(mesh_), v((size_t)n, 0.0);
        if (velocities_) indexer_->gatherFree(mesh_, velocities_, v.data());
        DeviceBuffer<Scalar> du, dv;
        if (du.upload(u.data(), n) || dv.upload(v.data(), n)) return;
        mat.deviceSync();
        Severe(od_plane_damping(mat.get(), (int)tets_.size(), planes_.data(), tets_.data(), coeff_, du.get(), dv.get()),
               "od_plane_damping");
        mat.deviceWritten();
    }

private:
    Scalar coeff_;
    Reference<Mesh> mesh_;
    Reference<NodeIndexer> indexer_;
    const Vector3 *velocities_;
    std::vector<Scalar> planes_;
    std::vector<int> tets_;
};

// ---- integrators (core/intergrator.h:13-35) ------------------------------------------------
class Intergrator {
public:
    Intergrator(int DOFS, Scalar massDamp, Scalar stiffDamp, Scalar ts)
        : dofs(DOFS), massDamping(massDamp), stiffnessDamping(stiffDamp), timeStep(ts) {}
    virtual void setExternalVirtualWork(const Forcer &forcer) = 0;
    virtual void addExternalDampingForceMatrix(const StaDampingForcer &forcer) {
        (void)forcer;
        std::fprintf(stderr, "Unimplemented features in Intergrator::setExternalDampingForceMatrix\n");
    }
    virtual void clearExternalDampingForceMatrix() {
        std::fprintf(stderr, "Unimplemented features in Intergrator::clearExternalDampingForceMatrix\n");
    }
    virtual void runOneTimeStep() = 0;
    virtual void updateMeshVerticesDisplacements(const Reference<NodeIndexer> &indexer, Reference<Mesh> &mesh) const = 0;
    virtual void getMeshVerticesVelocities(const Reference<NodeIndexer> &indexer, const Reference<Mesh> &mesh,
                                           Vector3 *velocities) const = 0;
    virtual ~Intergrator() = default;

protected:
    int dofs;
    Scalar massDamping;
    Scalar stiffnessDamping;
    Scalar timeStep;
};

// The device-resident integrators: every step runs on the device (element pass, assembly,
// factor, PCG, corrector); the linear solver given is a GpuCGSolver whose tolerance, iteration
// cap, preconditioner and precision the step's solve uses.
class GpuIntergrator : public Intergrator {
public:
    using Intergrator::Intergrator;
    ~GpuIntergrator() override { if (it_) od_integrator_destroy(it_); }
    void setExternalVirtualWork(const Forcer &forcer) override {
        std::vector<Scalar> vws((size_t)dofs);
        forcer.getVirtualWorks(dofs, dofs, nullptr, vws.data());
        status_ = Severe(od_integrator_set_external(it_, vws.data()), "od_integrator_set_external");
    }
    void runOneTimeStep() override { status_ = Severe(od_integrator_step(it_, 1), "od_integrator_step"); }
    void updateMeshVerticesDisplacements(const Reference<NodeIndexer> &indexer, Reference<Mesh> &mesh) const override {
        std::vector<Scalar> d((size_t)dofs);
        if (Severe(od_integrator_get_state(it_, d.data(), nullptr, nullptr), "od_integrator_get_state")) return;
        std::vector<Vector3> u((size_t)mesh->getNodeCount());
        indexer->scatterFree(mesh, d.data(), u.data());
        for (int n = 0; n < mesh->getNodeCount(); n++) mesh->getVertexDisplacement(n) = u[n];
    }
    void getMeshVerticesVelocities(const Reference<NodeIndexer> &indexer, const Reference<Mesh> &mesh,
                                   Vector3 *velocities) const override {
        std::vector<Scalar> v((size_t)dofs);
        if (Severe(od_integrator_get_state(it_, nullptr, v.data(), nullptr), "od_integrator_get_state")) return;
        indexer->scatterFree(mesh, v.data(), velocities);
    }
    // one-step nodal impulse (no reference counterpart): added to the next step's load only
    void setImpulse(const Forcer &forcer) {
        std::vector<Scalar> vws((size_t)dofs);
        forcer.getVirtualWorks(dofs, dofs, nullptr, vws.data());
        status_ = Severe(od_integrator_set_impulse(it_, vws.data()), "od_integrator_set_impulse");
    }
    int lastIterations() const {
        int it = 0;
        od_integrator_stats(it_, &it, nullptr, nullptr, nullptr);
        return it;
    }
    od_integrator *get() const { return it_; }
    int lastStatus() const { return status_; }

protected:
    int create(const Reference<Mesh> &mesh, const Reference<NodeIndexer> &indexer,
               const Reference<FullOrderNonlinearMaterial<BlockedSymSpMatrix>> &material,
               const Reference<LinearSolver<BlockedSymSpMatrix>> &linearSolver, bool newmark, Scalar beta, Scalar gamma) {
        const GpuCGSolver *cg = dynamic_cast<const GpuCGSolver *>(linearSolver.get());
        if (!cg) {
            std::fprintf(stderr, "odgpu: the device integrators take a GpuCGSolver\n");
            return OD_ERR_ARG;
        }
        mesh_ = mesh;
        indexer_ = indexer;
        int st = material->bindTo(mesh, indexer);
        if (st) return st;
        const Preconditioner &p = *cg->getPreconditioner();
        st = newmark ? Severe(od_newmark_create(indexer->deviceMesh(), beta, gamma, timeStep, massDamping, stiffnessDamping,
                                                p.deviceKind(), p.deviceParam1(), cg->tolerance(), cg->maxIterations(),
                                                cg->precision(), 1, &it_),
                              "od_newmark_create")
                     : Severe(od_backward_euler_create(indexer->deviceMesh(), timeStep, massDamping, stiffnessDamping,
                                                       p.deviceKind(), p.deviceParam1(), cg->tolerance(),
                                                       cg->maxIterations(), &it_),
                              "od_backward_euler_create");
        if (st) return st;
        od_solver *s = nullptr;
        st = Severe(od_integrator_solver(it_, &s), "od_integrator_solver");
        if (st == OD_OK) st = Severe(od_solver_set_precond_params(s, p.deviceParam1(), p.deviceParam2()),
                                     "od_solver_set_precond_params");
        return st;
    }
    Reference<Mesh> mesh_;
    Reference<NodeIndexer> indexer_;
    od_integrator *it_ = nullptr;
    int status_ = OD_OK;
};

// Full-order linearised Newmark (the reference's update, intergrator/linearNewmark.cpp:27-30,
// 57-79, at full order with the assembly structure of NonlinearImplicitBackwardEuler), with the
// constructor arguments of NonlinearImplicitBackwardEuler (nonlinearImplicitBackwardEuler.h:18-20)
// plus beta and gamma.
class GpuNewmark : public GpuIntergrator {
public:
    GpuNewmark(int DOFS, Scalar beta, Scalar gamma, Scalar massDamp, Scalar stiffDamp, Scalar ts,
               const Reference<Mesh> &m, const Reference<NodeIndexer> &nodeIndexer,
               const Reference<FullOrderNonlinearMaterial<BlockedSymSpMatrix>> &mater,
               const Reference<LinearSolver<BlockedSymSpMatrix>> &linearSolver)
        : GpuIntergrator(DOFS, massDamp, stiffDamp, ts) {
        status_ = create(m, nodeIndexer, mater, linearSolver, true, beta, gamma);
    }
    // core/intergrator.h:19-24: the base reports external damping unimplemented
    void addExternalDampingForceMatrix(const StaDampingForcer &forcer) override {
        Intergrator::addExternalDampingForceMatrix(forcer);
        status_ = OD_ERR_UNSUPPORTED;
    }
    void clearExternalDampingForceMatrix() override {
        Intergrator::clearExternalDampingForceMatrix();
        status_ = OD_ERR_UNSUPPORTED;
    }
};

// NonlinearImplicitBackwardEuler(DOFS, massDamp, stiffDamp, ts, mesh, indexer, material,
// linearSolver) (intergrator/nonlinearImplicitBackwardEuler.h:16-47).  updateMeshVerticesDisplacements
// first advances u += h v (nonlinearImplicitBackwardEuler.h:103-116).
class GpuBackwardEuler : public GpuIntergrator {
public:
    GpuBackwardEuler(int DOFS, Scalar massDamp, Scalar stiffDamp, Scalar ts, const Reference<Mesh> &m,
                     const Reference<NodeIndexer> &nodeIndexer,
                     const Reference<FullOrderNonlinearMaterial<BlockedSymSpMatrix>> &mater,
                     const Reference<LinearSolver<BlockedSymSpMatrix>> &linearSolver)
        : GpuIntergrator(DOFS, massDamp, stiffDamp, ts) {
        status_ = create(m, nodeIndexer, mater, linearSolver, false, 0.0, 0.0);
    }
    void updateMeshVerticesDisplacements(const Reference<NodeIndexer> &indexer, Reference<Mesh> &mesh) const override {
        if (Severe(od_backward_euler_update_mesh(it_), "od_backward_euler_update_mesh")) return;
        GpuIntergrator::updateMeshVerticesDisplacements(indexer, mesh);
    }
    // nonlinearImplicitBackwardEuler.h:122-128: the forcer adds into the damping matrix
    void addExternalDampingForceMatrix(const StaDampingForcer &forcer) override {
        od_matrix *D = nullptr;
        status_ = Severe(od_integrator_damping_matrix(it_, &D), "od_integrator_damping_matrix");
        if (status_) return;
        BlockedSymSpMatrix view(indexer_, D);
        SparseSymMatrixIndicesPerElementCache none;
        forcer.addDampingMatrix(none, view);
    }
    void clearExternalDampingForceMatrix() override {
        status_ = Severe(od_integrator_clear_damping(it_), "od_integrator_clear_damping");
    }
};

// ---- the driver (core/simulator.h:12-28, core/simulator.cpp:24-38) ---------------------------
class Simulator {
public:
    Simulator() : intergrator(nullptr) {}
    Simulator(const Reference<Mesh> m, const Reference<NodeIndexer> &nodeIndexer, Intergrator *intergrate)
        : mesh(m), indexer(nodeIndexer), intergrator(intergrate) {}
    Simulator(const Simulator &) = delete;
    Simulator &operator=(const Simulator &) = delete;
    void getVertexPositions(Vector3f *vertices) {
        intergrator->updateMeshVerticesDisplacements(indexer, mesh);
        for (int v = 0; v < mesh->getNodeCount(); v++) {
            const Vector3 p = mesh->getVertex(v) + mesh->getVertexDisplacement(v);
            for (int a = 0; a < 3; a++) vertices[v][a] = float(p[a]);
        }
    }
    void runTimeSteps(int stepCount) {
        for (int step = 0; step < stepCount; step++) intergrator->runOneTimeStep();
    }
    ~Simulator() { delete intergrator; }   // the simulator owns its integrator (core/simulator.cpp:41-43)

private:
    Reference<Mesh> mesh;
    Reference<NodeIndexer> indexer;
    Intergrator *intergrator;
};

// The reference's class names for the GPU implementations.
using CGSolver = GpuCGSolver;
using JacobiPreconditioner = GpuJacobiPreconditioner;
using InCholeskyPreconditioner = GpuInCholeskyPreconditioner;
using InLDLTPreconditioner = GpuInLDLTPreconditioner;
using CorotationalHyperelasticMaterial = GpuCorotationalMaterial;
using InvertibleNeoHookeanMaterial = GpuNeoHookeanMaterial;
using NonlinearImplicitBackwardEuler = GpuBackwardEuler;
using BlockedSymSpMatrixAssembler = BlockedSymSpMatrix::Assembler;

}  // namespace odgpu
