// Test shim (CPU only): builds the library's host topology (opendeformer_amd/csrc/od_setup.cpp,
// compiled with g++) and hands its arrays to the Python tests, which compare the layout with
// the oracle's and with a brute-force derivation.
#include <cstring>
#include <string>
#include <vector>

#include "../../opendeformer_amd/csrc/od_setup.h"

namespace {
od::Topology g_t;
std::string g_err;
template <typename T>
void put(const std::vector<T> &v, T *out) {
    if (out) std::memcpy(out, v.data(), sizeof(T) * v.size());
}
}  // namespace

extern "C" {
// sizes[12]: ndof (kernel vectors), ndofFree, nRows, nBlocks, nValues, nValuesRef, reference
// pattern columns, reference blockRows, vSrc entries, dSrc entries, partial, nLevels.
// Returns 0, or -1 with the error text in tp_error().
int tp_build(int nNodes, const double *X, int nTets, const int *tets, int nCon, const int *con, long long *sizes) {
    g_t = od::Topology();
    g_t.nNodes = nNodes;
    g_t.nTets = nTets;
    g_t.X.assign(X, X + 3 * (size_t)nNodes);
    g_t.tets.assign(tets, tets + 4 * (size_t)nTets);
    g_t.constrained.assign(con, con + nCon);
    g_err = od::buildTopology(g_t, 64);
    if (!g_err.empty()) return -1;
    const od::Topology &t = g_t;
    const bool p = t.partial;
    long long v[12] = {t.ndof, t.ndofFree, t.nRows, t.nBlocks, t.nValues, t.nValuesRef,
                       p ? (long long)t.refPcol.size() - 1 : t.nRows, p ? (long long)t.refRows.size() : t.nBlocks,
                       (long long)t.vSrc.size(), (long long)t.dSrc.size(), p ? 1 : 0, t.nLevels};
    std::memcpy(sizes, v, sizeof(v));
    return 0;
}
const char *tp_error(void) { return g_err.c_str(); }
// the reference's layout (blockPcol, blockColumnOris, blockRows) and the 78-entry map
void tp_ref_pattern(int *pcol, int *oris, int *rows, int *map78) {
    const od::Topology &t = g_t;
    if (t.partial) {
        put(t.refPcol, pcol); put(t.refOris, oris); put(t.refRows, rows);
    } else {
        put(t.rowPtr, pcol); put(t.valOff, oris);
        if (rows) for (size_t b = 0; b < t.colIdx.size(); b++) rows[b] = 3 * t.colIdx[b];
    }
    put(t.elemMap78, map78);
}
// the kernels' layout and the value-level lists (partial meshes)
void tp_internal(int *rowPtr, int *colIdx, int *valOff, int *blkOff, int *refPos, int *vPtr, int *vSrc, int *dPtr,
                 int *dSrc, int *nodeDof) {
    const od::Topology &t = g_t;
    put(t.rowPtr, rowPtr); put(t.colIdx, colIdx); put(t.valOff, valOff); put(t.blkOff, blkOff);
    put(t.refPos, refPos); put(t.vPtr, vPtr); put(t.vSrc, vSrc); put(t.dPtr, dPtr); put(t.dSrc, dSrc);
    put(t.nodeDof, nodeDof);
}
}
