// Drives the C++ host classes of od_gpu.hpp the way host code written against
// the reference's interfaces does: material -> generateMatrixAndVirtualWorks,
// SpMDV, CGSolver::solveLinearSystem, Intergrator::runOneTimeStep and
// updateMeshVerticesDisplacements, and a backward-Euler run with a plane collision
// damping forcer (addExternalDampingForceMatrix / clearExternalDampingForceMatrix).  Input and output are flat binary files
// written and read by tests/test_gpu_host_api.py, which compares every output
// with the oracle.
//
// usage: host_api_check <input.bin> <output.bin>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "od_gpu.hpp"

using odgpu::Scalar;

template <typename T> static bool rd(FILE *f, T *p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <typename T> static void wr(FILE *f, const T *p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) { std::perror("input"); return 2; }
    int hdr[5];  // nNodes, nTets, nConstrained, ndof, steps
    if (!rd(in, hdr, 5)) return 3;
    const int nNodes = hdr[0], nTets = hdr[1], nCon = hdr[2], ndof = hdr[3], steps = hdr[4];
    std::vector<Scalar> X(3 * (size_t)nNodes), par(4), uNode(3 * (size_t)nNodes), x(ndof), y(ndof), b(ndof), g(ndof);
    std::vector<int> tets(4 * (size_t)nTets), con(nCon);
    if (!rd(in, X.data(), X.size()) || !rd(in, tets.data(), tets.size()) || !rd(in, con.data(), con.size()) ||
        !rd(in, par.data(), 4) || !rd(in, uNode.data(), uNode.size()) || !rd(in, x.data(), ndof) ||
        !rd(in, y.data(), ndof) || !rd(in, b.data(), ndof) || !rd(in, g.data(), ndof))
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

    odgpu::Device dev(0);
    if (dev.lastStatus()) return 4;
    std::set<int> constrains(con.begin(), con.end());
    odgpu::TetMesh mesh(dev, nNodes, X.data(), nTets, tets.data(), constrains);
    if (mesh.lastStatus() || mesh.getMatrixOrder() != ndof) return 5;
    odgpu::CorotationalHyperelasticMaterial material(rho, E, nu, 0.0);

    // tangent stiffness and internal virtual works at the given displacements
    for (size_t i = 0; i < uNode.size(); i++) mesh.getDisplacements()[i] = uNode[i];
    odgpu::BlockedSymSpMatrix K(mesh);
    K.setZeros();
    std::vector<Scalar> vws(ndof, 0.0);
    if (material.generateMatrixAndVirtualWorks(mesh, K, vws.data())) return 6;
    std::vector<Scalar> Kv = K.hostValues();

    // SpMDV accumulates into y
    if (odgpu::SpMDV(K, x.data(), y.data())) return 7;

    // A = M + dt^2/4 K, solved by IC(0)-PCG from a zero initial guess
    odgpu::BlockedSymSpMatrix A(mesh);
    A.setZeros();
    if (material.generateMassMatrix(mesh, A)) return 8;
    A.Add(0.25 * dt * dt, K);
    odgpu::IC0Preconditioner ic0(1e-3);
    odgpu::CGSolver cg(1e-10, 5000, ic0, &A);
    std::vector<Scalar> sol(ndof, 0.0);
    cg.solveLinearSystem(b.data(), sol.data());
    if (cg.lastStatus()) return 9;
    int cgIters = cg.lastIterations();

    // Newmark from rest under the given external virtual works
    odgpu::NewmarkIntegrator nm(mesh, material, 0.25, 0.5, dt, 0.0, 0.0, 1e-6, 10000, ic0);
    if (nm.lastStatus()) return 10;
    nm.setExternalVirtualWork(g.data());
    for (int s = 0; s < steps; s++) {
        nm.runOneTimeStep();
        if (nm.lastStatus()) return 11;
    }
    nm.updateMeshVerticesDisplacements(mesh);
    std::vector<Scalar> vel(3 * (size_t)nNodes);
    nm.getMeshVerticesVelocities(mesh, vel.data());
    std::vector<Scalar> nmDisp(mesh.getDisplacements(), mesh.getDisplacements() + 3 * (size_t)nNodes);

    // backward Euler with a plane collision damping forcer: every step clears the damping
    // matrix, sets the pairs (setCollisionPair, then addCollisionPair) and adds it
    auto pairPlane = [&](int k) {
        odgpu::StaticPlaneShape pl;
        for (int a = 0; a < 3; a++) pl.planeNormal[a] = planes[4 * (size_t)k + a];
        pl.planeConstant = planes[4 * (size_t)k + 3];
        return pl;
    };
    odgpu::MeshPlaneCollisionDampingForcer forcer(coeff);
    nm.addExternalDampingForceMatrix(forcer);             // Newmark: reported unimplemented
    if (nm.lastStatus() != OD_ERR_UNSUPPORTED) return 12;
    odgpu::NonlinearImplicitBackwardEuler be(mesh, material, 0.0, 0.0, dt, 1e-6, 10000, ic0);
    if (be.lastStatus()) return 13;
    be.setExternalVirtualWork(g.data());
    for (int s = 0; s < steps; s++) {
        be.clearExternalDampingForceMatrix();
        forcer.setCollisionPair(pairPlane(0), pairTets[0]);
        for (int k = 1; k < nPairs; k++) forcer.addCollisionPair(pairPlane(k), pairTets[k]);
        be.addExternalDampingForceMatrix(forcer);
        be.runOneTimeStep();
        be.updateMeshVerticesDisplacements(mesh);
        if (be.lastStatus()) return 14;
    }
    std::vector<Scalar> beVel(3 * (size_t)nNodes);
    be.getMeshVerticesVelocities(mesh, beVel.data());

    FILE *out = std::fopen(argv[2], "wb");
    if (!out) { std::perror("output"); return 2; }
    long long nv = K.getValueCount();
    wr(out, &nv, 1);
    wr(out, Kv.data(), Kv.size());
    wr(out, vws.data(), vws.size());
    wr(out, y.data(), y.size());
    wr(out, sol.data(), sol.size());
    wr(out, &cgIters, 1);
    wr(out, nmDisp.data(), nmDisp.size());
    wr(out, vel.data(), vel.size());
    wr(out, beVel.data(), beVel.size());
    std::fclose(out);
    std::printf("host_api_check: ndof=%d values=%lld cg_iters=%d steps=%d\n", ndof, nv, cgIters, steps);
    return 0;
}
