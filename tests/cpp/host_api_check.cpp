// Drives the C++ host classes of od_gpu.hpp the way host code written against the reference's
// interfaces does: TetMesh + Constrainer + NodeIndexer, a material's getMatrixStructure,
// getPrecomputes, getMatrixIndicesPerElement and generateMatrixAndVirtualWorks, SpMDV, a
// CGSolver (LinearSolver) with an IC(0) Preconditioner, a Forcer's body force handed to
// Intergrator::setExternalVirtualWork, a Simulator stepping a GpuNewmark and reading vertex
// positions, and a backward-Euler run with a MeshPlaneCollisionDampingForcer
// (addExternalDampingForceMatrix / clearExternalDampingForceMatrix).  Input and output are flat
// binary files written and read by tests/test_gpu_host_api.py, which compares every output
// with the oracle.
//
// usage: host_api_check <input.bin> <output.bin>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "od_gpu.hpp"

using namespace odgpu;

template <typename T> static bool rd(FILE *f, T *p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <typename T> static void wr(FILE *f, const T *p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

static Reference<NodeIndexer> indexerFor(const Reference<Mesh> &mesh, std::vector<int> &fixedNodes) {
    Constrainer constrainer(mesh);
    constrainer.setConstrains((int)fixedNodes.size(), fixedNodes.data());
    return new NodeIndexer(constrainer);
}

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) { std::perror("input"); return 2; }
    int hdr[5];  // nNodes, nTets, nConstrained, ndof, steps
    if (!rd(in, hdr, 5)) return 3;
    const int nNodes = hdr[0], nTets = hdr[1], nCon = hdr[2], ndof = hdr[3], steps = hdr[4];
    std::vector<Scalar> X(3 * (size_t)nNodes), par(4), uNode(3 * (size_t)nNodes), x(ndof), y(ndof), b(ndof);
    std::vector<int> tets(4 * (size_t)nTets), con(nCon);
    if (!rd(in, X.data(), X.size()) || !rd(in, tets.data(), tets.size()) || !rd(in, con.data(), con.size()) ||
        !rd(in, par.data(), 4) || !rd(in, uNode.data(), uNode.size()) || !rd(in, x.data(), ndof) ||
        !rd(in, y.data(), ndof) || !rd(in, b.data(), ndof))
        return 3;
    // plane collision pairs for the backward-Euler run: count, 4 doubles per plane, tets, coefficient
    int nPairs = 0;
    if (!rd(in, &nPairs, 1) || nPairs < 1) return 3;
    std::vector<Scalar> planes(4 * (size_t)nPairs);
    std::vector<int> pairTets(nPairs);
    Scalar coeff = 0;
    if (!rd(in, planes.data(), planes.size()) || !rd(in, pairTets.data(), pairTets.size()) || !rd(in, &coeff, 1))
        return 3;
    std::fclose(in);
    const Scalar rho = par[0], E = par[1], nu = par[2], dt = par[3];
    if (Device::instance().lastStatus()) return 4;

    // mesh, whole-node constraints (Constrainer::setConstrains over the fixed nodes), indexer
    std::set<int> nodeSet;
    for (int d : con) nodeSet.insert(d / 3);
    std::vector<int> fixedNodes(nodeSet.begin(), nodeSet.end());
    Reference<Mesh> mesh = new TetMesh(nNodes, X.data(), nTets, tets.data());
    Reference<NodeIndexer> indexer = indexerFor(mesh, fixedNodes);
    if (indexer->getMatrixOrder(mesh) != ndof) return 5;
    Reference<FullOrderNonlinearMaterial<BlockedSymSpMatrix>> material =
        new CorotationalHyperelasticMaterial(rho, E, nu, 0.0);

    // the reference's set-up sequence: structure, matrix, precomputes, per-element indices
    BlockedSymSpMatrixAssembler assembler(ndof);
    material->getMatrixStructure(mesh, indexer, assembler);
    BlockedSymSpMatrix K(assembler);
    if (K.lastStatus() || K.getNumColumns() != ndof) return 5;
    FullOrderNonlinearMaterialCache cache = material->getPrecomputes(mesh);
    SparseSymMatrixIndicesPerElementCache indices = material->getMatrixIndicesPerElement(mesh, indexer, K);

    // tangent stiffness and internal virtual works at the given displacements
    for (int n = 0; n < nNodes; n++)
        mesh->getVertexDisplacement(n) = Vector3(uNode[3 * n], uNode[3 * n + 1], uNode[3 * n + 2]);
    K.setZeros();
    std::vector<Scalar> vws(ndof, 0.0);
    material->generateMatrixAndVirtualWorks(mesh, indexer, cache, indices, K, vws.data());
    if (K.lastStatus()) return 6;
    std::vector<Scalar> Kv = K.hostValues();

    // SpMDV accumulates into y
    if (SpMDV(K, x.data(), y.data())) return 7;

    // A = M + dt^2/4 K, solved by IC(0)-PCG from a zero initial guess
    BlockedSymSpMatrix A(assembler);
    A.setZeros();
    material->generateMassMatrix(mesh, indexer, indices, A);
    A.Add(0.25 * dt * dt, K);
    Reference<Preconditioner> ic0 = new GpuIC0Preconditioner(1e-3);
    Reference<LinearSolver<BlockedSymSpMatrix>> cg = new CGSolver(1e-10, 5000, ic0, &A);
    std::vector<Scalar> sol(ndof, 0.0);
    cg->solveLinearSystem(b.data(), sol.data());
    const GpuCGSolver &cgs = *static_cast<GpuCGSolver *>(cg.get());
    if (cgs.lastStatus()) return 9;
    int cgIters = cgs.lastIterations();
    // Preconditioner::Preprocess, then solvePreconditionerSystem (as CGSolver calls them)
    Reference<Preconditioner> threshold = new InCholeskyPreconditioner(1e-3, 1.0);
    std::vector<Scalar> zIc0(ndof), zIct(ndof);
    ic0->Preprocess(A);
    ic0->solvePreconditionerSystem(ndof, b.data(), zIc0.data());
    threshold->Preprocess(A);
    threshold->solvePreconditionerSystem(ndof, b.data(), zIct.data());

    // gravity through a Forcer, Newmark from rest stepped by a Simulator
    Scalar body[3] = {0.0, 0.0, -9.81 * rho};
    Forcer forcer;
    forcer.addBodyForce(mesh, body, indexer);
    std::vector<Scalar> g(ndof);
    forcer.getVirtualWorks(ndof, ndof, nullptr, g.data());
    Reference<LinearSolver<BlockedSymSpMatrix>> stepSolver = new CGSolver(1e-6, 10000, ic0);
    for (int n = 0; n < nNodes; n++) mesh->getVertexDisplacement(n) = Vector3();
    GpuNewmark *nm = new GpuNewmark(ndof, 0.25, 0.5, 0.0, 0.0, dt, mesh, indexer, material, stepSolver);
    if (nm->lastStatus()) return 10;
    nm->setExternalVirtualWork(forcer);
    std::vector<Vector3f> positions((size_t)nNodes);
    std::vector<Vector3> vel((size_t)nNodes);
    {
        Simulator sim(mesh, indexer, nm);   // owns nm from here
        sim.runTimeSteps(steps);
        if (nm->lastStatus()) return 11;
        sim.getVertexPositions(positions.data());
        nm->getMeshVerticesVelocities(indexer, mesh, vel.data());
        // Newmark: external damping reported unimplemented
        MeshPlaneCollisionDampingForcer none(coeff, mesh, indexer, vel.data());
        nm->addExternalDampingForceMatrix(none);
        if (nm->lastStatus() != OD_ERR_UNSUPPORTED) return 12;
    }
    std::vector<Scalar> nmDisp(3 * (size_t)nNodes);
    for (int n = 0; n < nNodes; n++)
        for (int a = 0; a < 3; a++) nmDisp[3 * (size_t)n + a] = mesh->getVertexDisplacement(n)[a];

    // backward Euler with a plane collision damping forcer on its own mesh: every step fills the
    // velocities, clears the damping matrix, sets the pairs and adds it, then steps
    Reference<Mesh> mesh2 = new TetMesh(nNodes, X.data(), nTets, tets.data());
    Reference<NodeIndexer> indexer2 = indexerFor(mesh2, fixedNodes);
    Reference<LinearSolver<BlockedSymSpMatrix>> beSolver = new CGSolver(1e-6, 10000, ic0);
    GpuBackwardEuler be(ndof, 0.0, 0.0, dt, mesh2, indexer2, material, beSolver);
    if (be.lastStatus()) return 13;
    be.setExternalVirtualWork(forcer);
    std::vector<Vector3> v2((size_t)nNodes);
    MeshPlaneCollisionDampingForcer damping(coeff, mesh2, indexer2, v2.data());
    auto pairPlane = [&](int k) {
        StaticPlaneShape pl;
        pl.planeNormal = Vector3(planes[4 * (size_t)k], planes[4 * (size_t)k + 1], planes[4 * (size_t)k + 2]);
        pl.planeConstant = planes[4 * (size_t)k + 3];
        return pl;
    };
    for (int s = 0; s < steps; s++) {
        be.getMeshVerticesVelocities(indexer2, mesh2, v2.data());
        be.clearExternalDampingForceMatrix();
        damping.setCollisionPair(pairPlane(0), pairTets[0]);
        for (int k = 1; k < nPairs; k++) damping.addCollisionPair(pairPlane(k), pairTets[k]);
        be.addExternalDampingForceMatrix(damping);
        be.runOneTimeStep();
        be.updateMeshVerticesDisplacements(indexer2, mesh2);
        if (be.lastStatus()) return 14;
    }
    std::vector<Vector3> beVel((size_t)nNodes);
    be.getMeshVerticesVelocities(indexer2, mesh2, beVel.data());

    FILE *out = std::fopen(argv[2], "wb");
    if (!out) { std::perror("output"); return 2; }
    long long nv = K.getNumEntries();
    wr(out, &nv, 1);
    wr(out, Kv.data(), Kv.size());
    wr(out, vws.data(), vws.size());
    wr(out, y.data(), y.size());
    wr(out, sol.data(), sol.size());
    wr(out, &cgIters, 1);
    wr(out, zIc0.data(), zIc0.size());
    wr(out, zIct.data(), zIct.size());
    wr(out, g.data(), g.size());
    wr(out, nmDisp.data(), nmDisp.size());
    for (const Vector3 &v : vel) wr(out, &v.x, 3);
    for (const Vector3f &p : positions) wr(out, &p.x, 3);
    for (const Vector3 &v : beVel) wr(out, &v.x, 3);
    std::fclose(out);
    std::printf("host_api_check: ndof=%d values=%lld cg_iters=%d steps=%d\n", ndof, nv, cgIters, steps);
    return 0;
}
