// od_gpu.hpp — C++ host classes over libodhip.so that keep the reference's
// solver, preconditioner, material and integrator interfaces, so that host
// code written against them switches to the GPU path by changing types only.
//
// Reference interfaces mirrored (paths relative to OpenDeformer/):
//   LinearSolver / CGSolver ............ core/linearSolver.h:12-20, solver/cgSolver.h:14-30
//   Preconditioner family .............. core/preconditioner.h:12-22,
//                                        solver/jacobiPreconditioner.h:10-22,
//                                        solver/inCholeskyPreconditioner.h:11-32
//   BlockedSymSpMatrix + SpMDV ......... core/sparseMatrix.h:186-295, 627-639;
//                                        core/numerMethod.h:134-136
//   FullOrderNonlinearMaterial ......... material/fullOrderNonlinear.h:36-44,
//                                        material/corotationalHyperelastic.h:11-17,
//                                        material/invertibleNeoHookean.h:11-16
//   Intergrator ........................ core/intergrator.h:13-35,
//                                        intergrator/nonlinearImplicitBackwardEuler.h:16-47
//   NodeIndexer ........................ core/nodeIndexer.cpp:8-74
//
// Semantics kept from the reference:
//   - host pointers in and out, DOF vectors over free DOFs;
//   - generateMatrixAndVirtualWorks and SpMDV accumulate into their outputs;
//   - solveLinearSystem treats `result` as the initial guess and stops when
//     ||r|| <= tol ||b|| or after maxIteration iterations;
//   - errors are reported, not thrown: the reference's Severe prints and
//     continues (core/oder.h:117-119); here Severe prints the library's error
//     text to stderr and the object records the failure in lastStatus().
// Differences: the matrix values live on the device and are mirrored to the
// host on demand (hostValues()/markHostDirty()); materials may differ per tet.
#pragma once
#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <vector>

#include "od.h"

namespace odgpu {

using Scalar = double;

inline int Severe(int status, const char *where) {
    if (status != OD_OK) std::fprintf(stderr, "odgpu: %s: %s\n", where, od_last_error());
    return status;
}

// Owns an od_ctx (one HIP stream on one device).
class Device {
public:
    explicit Device(int device = 0) { status_ = Severe(od_ctx_create(device, &ctx_), "od_ctx_create"); }
    ~Device() { if (ctx_) od_ctx_destroy(ctx_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    od_ctx *get() const { return ctx_; }
    int lastStatus() const { return status_; }

private:
    od_ctx *ctx_ = nullptr;
    int status_ = OD_OK;
};

// Device buffer of T, sized on first use.
template <typename T> class DeviceBuffer {
public:
    explicit DeviceBuffer(const Device &d) : dev_(d) {}
    ~DeviceBuffer() { if (p_) od_dev_free(dev_.get(), p_); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    T *ensure(long long n) {
        if (n <= n_) return p_;
        if (p_) od_dev_free(dev_.get(), p_);
        p_ = nullptr;
        n_ = 0;
        void *q = nullptr;
        if (Severe(od_dev_alloc(dev_.get(), n * (long long)sizeof(T), &q), "od_dev_alloc") == OD_OK) {
            p_ = (T *)q;
            n_ = n;
        }
        return p_;
    }
    int upload(const T *host, long long n) {
        if (!ensure(n)) return OD_ERR_HIP;
        return Severe(od_copy_h2d(dev_.get(), p_, host, n * (long long)sizeof(T)), "od_copy_h2d");
    }
    int download(T *host, long long n) const {
        return Severe(od_copy_d2h(dev_.get(), host, p_, n * (long long)sizeof(T)), "od_copy_d2h");
    }
    T *get() const { return p_; }

private:
    const Device &dev_;
    T *p_ = nullptr;
    long long n_ = 0;
};

// Tet mesh, its DOF indexing and its canonical matrix structure.  Replaces the
// set-up of Mesh + NodeIndexer + getMatrixStructure/getMatrixIndicesPerElement.
// `constrains` is Constrainer::outputConstrains' set of DOF ids (3*node+axis).
class TetMesh {
public:
    TetMesh(const Device &d, int nNodes, const Scalar *X, int nTets, const int *tets, const std::set<int> &constrains)
        : dev_(d), nNodes_(nNodes), nTets_(nTets), constrained_(constrains.begin(), constrains.end()),
          u_(3 * (size_t)nNodes, 0.0) {
        status_ = Severe(od_mesh_create(d.get(), nNodes, X, nTets, tets, (int)constrained_.size(),
                                        constrained_.data(), 1, nullptr, nullptr, &mesh_),
                         "od_mesh_create");
        if (status_ == OD_OK) {
            long long info[16] = {0};
            od_mesh_info(mesh_, info);
            order_ = (int)info[0];
        }
    }
    ~TetMesh() { if (mesh_) od_mesh_destroy(mesh_); }
    TetMesh(const TetMesh &) = delete;
    TetMesh &operator=(const TetMesh &) = delete;

    // NodeIndexer::getMatrixOrder / getGlobalIndex (core/nodeIndexer.cpp:8-25)
    int getMatrixOrder() const { return order_; }
    int getGlobalIndex(int node, int axis) const {
        const int dof = 3 * node + axis;
        size_t lo = 0, hi = constrained_.size();
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (constrained_[mid] < dof) lo = mid + 1; else hi = mid;
        }
        if (lo < constrained_.size() && constrained_[lo] == dof) return -1;
        return dof - (int)lo;
    }
    int getNodeCount() const { return nNodes_; }
    int getTetCount() const { return nTets_; }
    // Mesh displacement state (core/mesh.h:37-39): 3 per node.
    Scalar *getDisplacements() { return u_.data(); }
    const Scalar *getDisplacements() const { return u_.data(); }
    // free-DOF compaction of the node displacements
    void gatherFree(const Scalar *perNode, Scalar *freeDofs) const {
        for (int n = 0; n < nNodes_; n++)
            for (int a = 0; a < 3; a++) {
                int g = getGlobalIndex(n, a);
                if (g >= 0) freeDofs[g] = perNode[3 * n + a];
            }
    }
    void scatterFree(const Scalar *freeDofs, Scalar *perNode) const {
        for (int n = 0; n < nNodes_; n++)
            for (int a = 0; a < 3; a++) {
                int g = getGlobalIndex(n, a);
                perNode[3 * n + a] = g >= 0 ? freeDofs[g] : 0.0;
            }
    }
    od_mesh *get() const { return mesh_; }
    const Device &device() const { return dev_; }
    int lastStatus() const { return status_; }

private:
    const Device &dev_;
    int nNodes_, nTets_;
    std::vector<int> constrained_;
    std::vector<Scalar> u_;
    od_mesh *mesh_ = nullptr;
    int order_ = 0;
    int status_ = OD_OK;
};

// BlockedSymSparseMatrix<3,3> over the canonical layout of its mesh.  Values
// live on the device; hostValues() mirrors them, markHostDirty() pushes host
// edits back before the next device use.
class BlockedSymSpMatrix {
public:
    explicit BlockedSymSpMatrix(const TetMesh &mesh) : mesh_(mesh) {
        status_ = Severe(od_matrix_create(mesh.get(), &m_), "od_matrix_create");
        long long info[16] = {0};
        od_mesh_info(mesh.get(), info);
        nValues_ = info[3];
    }
    ~BlockedSymSpMatrix() { if (m_) od_matrix_destroy(m_); }
    BlockedSymSpMatrix(const BlockedSymSpMatrix &) = delete;
    BlockedSymSpMatrix &operator=(const BlockedSymSpMatrix &) = delete;

    int getNumColumns() const { return mesh_.getMatrixOrder(); }
    int getNumRows() const { return mesh_.getMatrixOrder(); }
    long long getValueCount() const { return nValues_; }
    // sparseMatrix.h:277-295
    void setZeros() { host_.clear(); hostDirty_ = false; status_ = Severe(od_matrix_zero(m_), "od_matrix_zero"); }
    void Scale(Scalar s) { deviceSync(); status_ = Severe(od_matrix_axpby(m_, s, 0.0, nullptr), "od_matrix_axpby"); host_.clear(); }
    void Add(Scalar s, const BlockedSymSpMatrix &o) {
        deviceSync();
        o.deviceSync();
        status_ = Severe(od_matrix_axpby(m_, 1.0, s, o.m_), "od_matrix_axpby");
        host_.clear();
    }
    // host mirror of `values` (core/sparseMatrix.h:627-635 layout)
    std::vector<Scalar> &hostValues() {
        if (host_.empty()) {
            host_.assign((size_t)nValues_, 0.0);
            status_ = Severe(od_matrix_download(m_, host_.data()), "od_matrix_download");
        }
        return host_;
    }
    void markHostDirty() { hostDirty_ = true; }
    // the canonical index arrays blockPcol / blockColumnOris / blockRows
    void getStructure(std::vector<int> &blockPcol, std::vector<int> &blockColumnOris, std::vector<int> &blockRows) const {
        long long info[16] = {0};
        od_mesh_info(mesh_.get(), info);
        blockPcol.resize((size_t)info[1] + 1);
        blockColumnOris.resize((size_t)info[1] + 1);
        blockRows.resize((size_t)info[2]);
        Severe(od_pattern_export(mesh_.get(), blockPcol.data(), blockColumnOris.data(), blockRows.data(), nullptr),
               "od_pattern_export");
    }
    // pushes host edits; invalidates the mirror after device writes
    void deviceSync() const {
        if (hostDirty_ && !host_.empty()) {
            Severe(od_matrix_upload(m_, host_.data()), "od_matrix_upload");
            hostDirty_ = false;
        }
    }
    void deviceWritten() { host_.clear(); hostDirty_ = false; }
    od_matrix *get() const { return m_; }
    const TetMesh &mesh() const { return mesh_; }
    int lastStatus() const { return status_; }

private:
    const TetMesh &mesh_;
    od_matrix *m_ = nullptr;
    long long nValues_ = 0;
    mutable std::vector<Scalar> host_;
    mutable bool hostDirty_ = false;
    int status_ = OD_OK;
};

// SpMDV (core/numerMethod.h:134-136): dest += mat * src, host vectors.
inline int SpMDV(const BlockedSymSpMatrix &mat, const Scalar *src, Scalar *dest) {
    const int n = mat.getNumColumns();
    const Device &d = mat.mesh().device();
    DeviceBuffer<Scalar> x(d), y(d);
    mat.deviceSync();
    int st = x.upload(src, n);
    if (st == OD_OK) st = y.upload(dest, n);
    if (st == OD_OK) st = Severe(od_spmv(mat.get(), x.get(), y.get(), 1), "od_spmv");
    if (st == OD_OK) st = y.download(dest, n);
    return st;
}

// Preconditioner family (core/preconditioner.h:12-22).  The identity base
// matches the reference's default Preconditioner.
class Preconditioner {
public:
    explicit Preconditioner(int kind = OD_PRE_NONE, Scalar param = 0.0) : kind_(kind), param_(param) {}
    virtual ~Preconditioner() { if (s_) od_solver_destroy(s_); }
    Preconditioner(const Preconditioner &) = delete;
    Preconditioner &operator=(const Preconditioner &) = delete;
    // Preprocess / resetPreconditionerSystem: the factor is rebuilt from the
    // bound matrix on every apply and inside every PCG solve, as the
    // reference's CGSolver calls Preprocess per solve (solver/cgSolver.cpp:23).
    virtual void resetPreconditionerSystem(const BlockedSymSpMatrix &mat) { mat_ = &mat; }
    virtual void Preprocess(const BlockedSymSpMatrix &mat) { mat_ = &mat; }
    virtual void solvePreconditionerSystem(int width, const Scalar *rhs, Scalar *result) const {
        if (!mat_ || kind_ == OD_PRE_NONE) { std::memcpy(result, rhs, sizeof(Scalar) * (size_t)width); return; }
        const Device &d = mat_->mesh().device();
        if (!s_ && Severe(od_solver_create(mat_->mesh().get(), kind_, param_, OD_FP64, &s_), "od_solver_create"))
            return;
        DeviceBuffer<Scalar> r(d), z(d);
        mat_->deviceSync();
        if (r.upload(rhs, width) || !z.ensure(width)) return;
        if (Severe(od_precond_apply(s_, mat_->get(), r.get(), z.get()), "od_precond_apply")) return;
        z.download(result, width);
    }
    int kind() const { return kind_; }
    Scalar param() const { return param_; }

private:
    int kind_;
    Scalar param_;
    const BlockedSymSpMatrix *mat_ = nullptr;
    mutable od_solver *s_ = nullptr;
};

// JacobiPreconditioner (solver/jacobiPreconditioner.h:10-22)
class JacobiPreconditioner : public Preconditioner {
public:
    JacobiPreconditioner() : Preconditioner(OD_PRE_JACOBI) {}
};

// IC(0) on the block pattern in the role of InCholeskyPreconditioner
// (solver/inCholeskyPreconditioner.h:11-32); `epsilon` is the pivot guard
// d <= 0 -> epsilon^2 d^2 of inCholeskyPreconditioner.cpp:52.
class IC0Preconditioner : public Preconditioner {
public:
    explicit IC0Preconditioner(Scalar epsilon = 1e-8) : Preconditioner(OD_PRE_IC0, epsilon) {}
};

// CGSolver (solver/cgSolver.h:14-30): same constructor order (tol, maxIter,
// preconditioner, matrix).  `precision` selects the FP32-inner form.
class CGSolver {
public:
    CGSolver(Scalar tol, int maxIter, const Preconditioner &precondition = Preconditioner(),
             const BlockedSymSpMatrix *m = nullptr, int precision = OD_FP64)
        : tol_(tol), maxIt_(maxIter), pkind_(precondition.kind()), pparam_(precondition.param()),
          precision_(precision), mat_(m) {}
    ~CGSolver() {
        b_.reset();
        x_.reset();
        if (s_) od_solver_destroy(s_);
    }
    CGSolver(const CGSolver &) = delete;
    CGSolver &operator=(const CGSolver &) = delete;
    void resetLinearSystem(const BlockedSymSpMatrix *m) { mat_ = m; }
    // result: initial guess in, solution out (solver/cgSolver.cpp:13-57)
    void solveLinearSystem(const Scalar *rhs, Scalar *result) {
        if (!mat_) { status_ = OD_ERR_STATE; std::fprintf(stderr, "odgpu: CGSolver without a matrix\n"); return; }
        const Device &d = mat_->mesh().device();
        const int n = mat_->getNumColumns();
        if (!s_) {
            status_ = Severe(od_solver_create(mat_->mesh().get(), pkind_, pparam_, precision_, &s_), "od_solver_create");
            if (status_) return;
        }
        if (!b_) { b_.reset(new DeviceBuffer<Scalar>(d)); x_.reset(new DeviceBuffer<Scalar>(d)); }
        mat_->deviceSync();
        if ((status_ = b_->upload(rhs, n)) || (status_ = x_->upload(result, n))) return;
        status_ = Severe(od_pcg_solve(s_, mat_->get(), b_->get(), x_->get(), tol_, maxIt_, &iters_, &relres_),
                         "od_pcg_solve");
        if (status_ == OD_OK) status_ = x_->download(result, n);
    }
    int lastIterations() const { return iters_; }
    Scalar lastRelativeResidual() const { return relres_; }
    int lastStatus() const { return status_; }

private:
    Scalar tol_;
    int maxIt_, pkind_;
    Scalar pparam_;
    int precision_;
    const BlockedSymSpMatrix *mat_;
    od_solver *s_ = nullptr;
    std::unique_ptr<DeviceBuffer<Scalar>> b_, x_;
    int iters_ = 0;
    Scalar relres_ = 0;
    int status_ = OD_OK;
};

// FullOrderNonlinearMaterial family (material/fullOrderNonlinear.h:36-44).
// The device stores (kind, E, nu, rho) per tet; a material object writes its
// parameters for every tet of the mesh it is used with.
class FullOrderNonlinearMaterial {
public:
    virtual ~FullOrderNonlinearMaterial() = default;
    // per-tet tables written to the mesh
    virtual void tables(int nTets, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                        std::vector<Scalar> &rho) const = 0;
    virtual Scalar corotationalThreshold() const { return 0.0; }
    virtual Scalar inversionThreshold() const { return 0.1; }
    int bindTo(const TetMesh &mesh) const {
        std::vector<int> kind;
        std::vector<Scalar> E, nu, rho;
        tables(mesh.getTetCount(), kind, E, nu, rho);
        return Severe(od_mesh_set_material(mesh.get(), kind.data(), E.data(), nu.data(), rho.data(),
                                           corotationalThreshold(), inversionThreshold()),
                      "od_mesh_set_material");
    }
    // generateMatrixAndVirtualWorks (fullOrderNonlinear.h:40-41): matrix +=
    // K(u) and vws += f_int(u) at the mesh's current node displacements.
    int generateMatrixAndVirtualWorks(const TetMesh &mesh, BlockedSymSpMatrix &matrix, Scalar *vws) const {
        int st = bindTo(mesh);
        if (st) return st;
        const int n = mesh.getMatrixOrder();
        std::vector<Scalar> u((size_t)n, 0.0), f((size_t)n, 0.0);
        mesh.gatherFree(mesh.getDisplacements(), u.data());
        DeviceBuffer<Scalar> du(mesh.device()), df(mesh.device());
        if ((st = du.upload(u.data(), n)) || (st = df.upload(f.data(), n))) return st;
        matrix.deviceSync();
        st = Severe(od_assemble(mesh.get(), du.get(), matrix.get(), df.get()), "od_assemble");
        matrix.deviceWritten();
        if (st == OD_OK) st = df.download(f.data(), n);
        if (st == OD_OK)
            for (int i = 0; i < n; i++) vws[i] += f[(size_t)i];
        return st;
    }
    // MechMaterial::generateMassMatrix (core/mechMaterial.cpp:69-98): M += mass
    int generateMassMatrix(const TetMesh &mesh, BlockedSymSpMatrix &M) const {
        int st = bindTo(mesh);
        if (st) return st;
        M.deviceSync();
        st = Severe(od_mass_matrix(mesh.get(), M.get()), "od_mass_matrix");
        M.deviceWritten();
        return st;
    }
};

// CorotationalHyperelasticMaterial(rho, youngs, possion, compressingTreshold)
// (material/corotationalHyperelastic.h:13)
class CorotationalHyperelasticMaterial : public FullOrderNonlinearMaterial {
public:
    CorotationalHyperelasticMaterial(Scalar rho, Scalar youngs, Scalar possion, Scalar compressingThreshold)
        : rho_(rho), E_(youngs), nu_(possion), thr_(compressingThreshold) {}
    void tables(int n, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                std::vector<Scalar> &rho) const override {
        kind.assign((size_t)n, OD_MAT_COROTATIONAL);
        E.assign((size_t)n, E_);
        nu.assign((size_t)n, nu_);
        rho.assign((size_t)n, rho_);
    }
    Scalar corotationalThreshold() const override { return thr_; }

private:
    Scalar rho_, E_, nu_, thr_;
};

// InvertibleNeoHookeanMaterial(rho, youngs, possion, inversionTrashold)
// (material/invertibleNeoHookean.h:13)
class InvertibleNeoHookeanMaterial : public FullOrderNonlinearMaterial {
public:
    InvertibleNeoHookeanMaterial(Scalar rho, Scalar youngs, Scalar possion, Scalar inversionThreshold)
        : rho_(rho), E_(youngs), nu_(possion), thr_(inversionThreshold) {}
    void tables(int n, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                std::vector<Scalar> &rho) const override {
        kind.assign((size_t)n, OD_MAT_NEOHOOKEAN);
        E.assign((size_t)n, E_);
        nu.assign((size_t)n, nu_);
        rho.assign((size_t)n, rho_);
    }
    Scalar inversionThreshold() const override { return thr_; }

private:
    Scalar rho_, E_, nu_, thr_;
};

// Per-tet material table (no reference counterpart: the reference binds one
// material per mesh).
class PerTetMaterial : public FullOrderNonlinearMaterial {
public:
    PerTetMaterial(std::vector<int> kind, std::vector<Scalar> E, std::vector<Scalar> nu, std::vector<Scalar> rho,
                   Scalar corotThreshold = 0.0, Scalar neoThreshold = 0.1)
        : kind_(std::move(kind)), E_(std::move(E)), nu_(std::move(nu)), rho_(std::move(rho)), ct_(corotThreshold),
          nt_(neoThreshold) {}
    void tables(int, std::vector<int> &kind, std::vector<Scalar> &E, std::vector<Scalar> &nu,
                std::vector<Scalar> &rho) const override {
        kind = kind_;
        E = E_;
        nu = nu_;
        rho = rho_;
    }
    Scalar corotationalThreshold() const override { return ct_; }
    Scalar inversionThreshold() const override { return nt_; }

private:
    std::vector<int> kind_;
    std::vector<Scalar> E_, nu_, rho_;
    Scalar ct_, nt_;
};

// Intergrator (core/intergrator.h:13-35).  setExternalVirtualWork takes the
// free-DOF virtual works that Forcer::getVirtualWorks produces
// (core/forcer.cpp:80-96); velocities come out as 3 per node.
// collision/collisionShapes.h:17-20: the plane n.x + c = 0 (below: n.x + c < 0)
struct StaticPlaneShape {
    Scalar planeNormal[3];
    Scalar planeConstant;
};

// MeshPlaneCollisionDampingForcer (collision/penaltyDampingForcer.h:25-38): the (plane, tet)
// pair whose damping matrix addExternalDampingForceMatrix adds.  The positions and velocities
// it reads are the integrator's own (the reference passes the mesh and a velocities array
// the caller fills from getMeshVerticesVelocities).  setCollisionPair replaces the pair as
// the reference's does; addCollisionPair queues further pairs so that one
// addExternalDampingForceMatrix call adds them all, in order.
class MeshPlaneCollisionDampingForcer {
public:
    explicit MeshPlaneCollisionDampingForcer(Scalar dampingCoefficient) : coeff_(dampingCoefficient) {}
    void setCollisionPair(const StaticPlaneShape &plane, int elementIndex) {
        planes_.clear();
        tets_.clear();
        addCollisionPair(plane, elementIndex);
    }
    void addCollisionPair(const StaticPlaneShape &plane, int elementIndex) {
        planes_.insert(planes_.end(), {plane.planeNormal[0], plane.planeNormal[1], plane.planeNormal[2],
                                       plane.planeConstant});
        tets_.push_back(elementIndex);
    }
    int pairCount() const { return (int)tets_.size(); }
    const Scalar *planes() const { return planes_.data(); }
    const int *elements() const { return tets_.data(); }
    Scalar coefficient() const { return coeff_; }

private:
    Scalar coeff_;
    std::vector<Scalar> planes_;
    std::vector<int> tets_;
};

class Intergrator {
public:
    Intergrator(int dofs, Scalar massDamp, Scalar stiffDamp, Scalar ts)
        : dofs(dofs), massDamping(massDamp), stiffnessDamping(stiffDamp), timeStep(ts) {}
    virtual ~Intergrator() { if (it_) od_integrator_destroy(it_); }
    Intergrator(const Intergrator &) = delete;
    Intergrator &operator=(const Intergrator &) = delete;

    virtual void setExternalVirtualWork(const Scalar *vws) {
        status_ = Severe(od_integrator_set_external(it_, vws), "od_integrator_set_external");
    }
    virtual void runOneTimeStep() { status_ = Severe(od_integrator_step(it_, 1), "od_integrator_step"); }
    virtual void updateMeshVerticesDisplacements(TetMesh &mesh) const {
        std::vector<Scalar> d((size_t)dofs);
        if (Severe(od_integrator_get_state(it_, d.data(), nullptr, nullptr), "od_integrator_get_state")) return;
        mesh.scatterFree(d.data(), mesh.getDisplacements());
    }
    virtual void getMeshVerticesVelocities(const TetMesh &mesh, Scalar *velocities) const {
        std::vector<Scalar> v((size_t)dofs);
        if (Severe(od_integrator_get_state(it_, nullptr, v.data(), nullptr), "od_integrator_get_state")) return;
        mesh.scatterFree(v.data(), velocities);
    }
    // core/intergrator.h:19-24: integrators without external damping report it unimplemented
    virtual void addExternalDampingForceMatrix(const MeshPlaneCollisionDampingForcer &) {
        std::fprintf(stderr, "Unimplemented features in Intergrator::addExternalDampingForceMatrix\n");
        status_ = OD_ERR_UNSUPPORTED;
    }
    virtual void clearExternalDampingForceMatrix() {
        std::fprintf(stderr, "Unimplemented features in Intergrator::clearExternalDampingForceMatrix\n");
        status_ = OD_ERR_UNSUPPORTED;
    }
    // one-step nodal impulse (added to the next step's external force only)
    void setImpulse(const Scalar *vws) { status_ = Severe(od_integrator_set_impulse(it_, vws), "od_integrator_set_impulse"); }
    int lastIterations() const {
        int it = 0;
        od_integrator_stats(it_, &it, nullptr, nullptr, nullptr);
        return it;
    }
    od_integrator *get() const { return it_; }
    int lastStatus() const { return status_; }

protected:
    int dofs;
    Scalar massDamping, stiffnessDamping, timeStep;
    od_integrator *it_ = nullptr;
    int status_ = OD_OK;
};

// Full-order linearised Newmark: the reference's Newmark update
// (intergrator/linearNewmark.cpp:27-30, 57-79) at full order with the
// assembly structure of NonlinearImplicitBackwardEuler.
class NewmarkIntegrator : public Intergrator {
public:
    NewmarkIntegrator(const TetMesh &mesh, const FullOrderNonlinearMaterial &material, Scalar beta, Scalar gamma,
                      Scalar ts, Scalar massDamp, Scalar stiffDamp, Scalar tol, int maxIter,
                      const Preconditioner &precondition, int precision = OD_FP64, bool initialAcceleration = true)
        : Intergrator(mesh.getMatrixOrder(), massDamp, stiffDamp, ts) {
        status_ = material.bindTo(mesh);
        if (status_ == OD_OK)
            status_ = Severe(od_newmark_create(mesh.get(), beta, gamma, ts, massDamp, stiffDamp, precondition.kind(),
                                               precondition.param(), tol, maxIter, precision,
                                               initialAcceleration ? 1 : 0, &it_),
                             "od_newmark_create");
    }
};

// NonlinearImplicitBackwardEuler (intergrator/nonlinearImplicitBackwardEuler.h:16-47):
// same constructor parameters, with the linear solver given as (tol, maxIter,
// preconditioner).  updateMeshVerticesDisplacements advances u += h v first,
// as the reference's does (nonlinearImplicitBackwardEuler.h:103-116).
class NonlinearImplicitBackwardEuler : public Intergrator {
public:
    NonlinearImplicitBackwardEuler(const TetMesh &mesh, const FullOrderNonlinearMaterial &material, Scalar massDamp,
                                   Scalar stiffDamp, Scalar ts, Scalar tol, int maxIter,
                                   const Preconditioner &precondition)
        : Intergrator(mesh.getMatrixOrder(), massDamp, stiffDamp, ts) {
        status_ = material.bindTo(mesh);
        if (status_ == OD_OK)
            status_ = Severe(od_backward_euler_create(mesh.get(), ts, massDamp, stiffDamp, precondition.kind(),
                                                      precondition.param(), tol, maxIter, &it_),
                             "od_backward_euler_create");
    }
    void updateMeshVerticesDisplacements(TetMesh &mesh) const override {
        if (Severe(od_backward_euler_update_mesh(it_), "od_backward_euler_update_mesh")) return;
        Intergrator::updateMeshVerticesDisplacements(mesh);
    }
    // nonlinearImplicitBackwardEuler.h:122-128
    void addExternalDampingForceMatrix(const MeshPlaneCollisionDampingForcer &forcer) override {
        status_ = Severe(od_integrator_add_plane_damping(it_, forcer.pairCount(), forcer.planes(), forcer.elements(),
                                                         forcer.coefficient()),
                         "od_integrator_add_plane_damping");
    }
    void clearExternalDampingForceMatrix() override {
        status_ = Severe(od_integrator_clear_damping(it_), "od_integrator_clear_damping");
    }
};

}  // namespace odgpu
