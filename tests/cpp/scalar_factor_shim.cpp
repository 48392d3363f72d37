// Test shim (CPU only): builds the library's host topology and scalar factorisations
// (opendeformer_amd/csrc/od_setup.cpp, od_precond_host.cpp, compiled with g++) and hands the
// factor and its sweep schedule to the Python tests, which compare them with the oracle.
#include <cstring>
#include <vector>

#include "../../opendeformer_amd/csrc/od_precond.h"
#include "../../opendeformer_amd/csrc/od_setup.h"

namespace {
od::ScalarFactor g_factor;
od::ScalarSweepsHost g_sweeps;
template <typename T>
void put(const std::vector<T> &v, T *out) {
    if (out) std::memcpy(out, v.data(), sizeof(T) * v.size());
}
}  // namespace

extern "C" {
// kind 2: threshold IC (p1 epsilon, p2 relaxedScale); 4: SAINV LDL^T (p1 sainv, p2 ldlt).
// sizes[10]: n, factor entries, forward entries, backward entries, forward levels,
// backward levels, forward launches, backward launches, 0, 0.  Returns 0 or -1.
int sf_build(int nNodes, const double *X, int nTets, const int *tets, int nCon, const int *con, const double *vals,
             int kind, double p1, double p2, long long *sizes) {
    od::Topology t;
    t.nNodes = nNodes;
    t.nTets = nTets;
    t.X.assign(X, X + 3 * (size_t)nNodes);
    t.tets.assign(tets, tets + 4 * (size_t)nTets);
    t.constrained.assign(con, con + nCon);
    if (!od::buildTopology(t, 64).empty()) return -1;
    if (kind == 2) od::thresholdIcFactor(t, vals, p1, p2, g_factor);
    else if (kind == 4) od::sainvLdltFactor(t, vals, p1, p2, g_factor);
    else return -1;
    od::buildScalarSweeps(g_factor, kind == 4, g_sweeps);
    const od::ScalarSweepsHost &s = g_sweeps;
    long long v[10] = {s.n, (long long)g_factor.values.size(), (long long)s.fCol.size(), (long long)s.bRow.size(),
                       (long long)s.fwd.levelPtr.size() - 1, (long long)s.bwd.levelPtr.size() - 1,
                       (long long)s.fwd.launches.size(), (long long)s.bwd.launches.size(), 0, 0};
    std::memcpy(sizes, v, sizeof(v));
    return 0;
}
void sf_factor(int *pcol, int *rows, double *values, double *invDiag) {
    put(g_factor.pcol, pcol);
    put(g_factor.rows, rows);
    put(g_factor.values, values);
    put(g_factor.invDiag, invDiag);
}
void sf_sweeps(int *fPtr, int *fCol, double *fVal, int *bPtr, int *bRow, double *bVal, double *diag, int *fLevelPtr,
               int *fLevelRows, int *bLevelPtr, int *bLevelRows) {
    const od::ScalarSweepsHost &s = g_sweeps;
    put(s.fPtr, fPtr); put(s.fCol, fCol); put(s.fVal, fVal);
    put(s.bPtr, bPtr); put(s.bRow, bRow); put(s.bVal, bVal); put(s.diag, diag);
    put(s.fwd.levelPtr, fLevelPtr); put(s.fwd.levelRows, fLevelRows);
    put(s.bwd.levelPtr, bLevelPtr); put(s.bwd.levelRows, bLevelRows);
}
}
