// Test shim (CPU only): builds the library's host topology and grouped-product layout
// (opendeformer_amd/csrc/od_setup.cpp, od_groups.cpp, compiled with g++) and replays, on the
// host, the data flow of the two device products on given values and x:
//   the slot product: every row's 16 slots (diagonal, upper blocks, lower blocks in
//     transposed-list order) as 0 + (e0 x0 + e1 x1 + e2 x2) with fused multiply-adds, summed
//     by the 16-lane tree (pairs at distance 8, 4, 2, 1);
//   the grouped product: the upper streams, LDS deposits and crossing lists of the layout,
//     then the same per-row slot sum.
// The Python test compares the two bit for bit, with a dense product, and checks that every
// deposit of every group is written exactly once.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../opendeformer_amd/csrc/od_setup.h"

#pragma STDC FP_CONTRACT OFF

namespace {
od::Topology g_t;
od::GroupLayout g_g;
std::string g_err;

struct V3 { double a, b, c; };
inline double row3(double e0, double e1, double e2, const double *x) {
    return std::fma(e2, x[2], std::fma(e1, x[1], e0 * x[0]));
}
// 0 + the block times x: upper (row-major e), transposed, or the packed diagonal
inline V3 prod(const double *e, const double *x, int kind) {
    double r0, r1, r2;
    if (kind == 0) { r0 = row3(e[0], e[1], e[2], x); r1 = row3(e[3], e[4], e[5], x); r2 = row3(e[6], e[7], e[8], x); }
    else if (kind == 1) { r0 = row3(e[0], e[3], e[6], x); r1 = row3(e[1], e[4], e[7], x); r2 = row3(e[2], e[5], e[8], x); }
    else { r0 = row3(e[0], e[1], e[2], x); r1 = row3(e[1], e[3], e[4], x); r2 = row3(e[2], e[4], e[5], x); }
    volatile double z = 0.0;
    return {z + r0, z + r1, z + r2};
}
inline V3 tree16(V3 v[16]) {
    for (int o = 8; o > 0; o >>= 1)
        for (int s = 0; s < o; s++) { v[s].a = v[s].a + v[s + o].a; v[s].b = v[s].b + v[s + o].b; v[s].c = v[s].c + v[s + o].c; }
    return v[0];
}
}  // namespace

extern "C" {
// stats[8]: ok, nGroups, maxDeps, nUp, nX, internal upper blocks, crossing upper blocks, nRows
int gs_build(int nNodes, const double *X, int nTets, const int *tets, int nCon, const int *con, long long *stats) {
    g_t = od::Topology();
    g_t.nNodes = nNodes;
    g_t.nTets = nTets;
    g_t.X.assign(X, X + 3 * (size_t)nNodes);
    g_t.tets.assign(tets, tets + 4 * (size_t)nTets);
    g_t.constrained.assign(con, con + nCon);
    g_err = od::buildTopology(g_t, 64);
    if (!g_err.empty()) return -1;
    g_err = od::buildGroupLayout(g_t, 64, 8, g_g);
    if (!g_err.empty()) return -1;
    long long in = 0, cr = 0;
    for (int m : g_g.upMeta) { if (m == -1) continue; if (m < 0) in++; else cr++; }
    long long v[8] = {g_g.ok ? 1 : 0, g_g.nGroups, g_g.maxDeps, g_g.nUp, g_g.nX, in, cr, g_t.nRows};
    std::memcpy(stats, v, sizeof(v));
    return 0;
}
const char *gs_error(void) { return g_err.c_str(); }
int gs_max_halo(void) { return g_g.maxHalo; }
int gs_nvalues(void) { return (int)g_t.nValues; }
// y_slot and y_grp (3 nRows each) from values (nValues, the canonical layout) and x (3 nRows).
// Returns the number of deposits not written exactly once plus the halo slots out of range, out
// of order or naming a row of the group (0 when the layout is complete).
long long gs_products(const double *vals, const double *x, double *ySlot, double *yGrp) {
    const od::Topology &t = g_t;
    for (int r = 0; r < t.nRows; r++) {
        V3 v[16];
        for (auto &q : v) q = {0, 0, 0};
        int s = 0;
        v[s++] = prod(vals + t.valOff[r], x + 3 * r, 2);
        for (int b = t.rowPtr[r] + 1; b < t.rowPtr[r + 1]; b++) v[s++] = prod(vals + t.blkOff[b], x + 3 * t.colIdx[b], 0);
        for (int k = t.tPtr[r]; k < t.tPtr[r + 1]; k++) v[s++] = prod(vals + t.blkOff[t.tBlk[k]], x + 3 * t.tRow[k], 1);
        const V3 y = tree16(v);
        ySlot[3 * r] = y.a; ySlot[3 * r + 1] = y.b; ySlot[3 * r + 2] = y.c;
    }
    if (!g_g.ok) return -1;
    const od::GroupLayout &g = g_g;
    long long bad = 0;
    for (int p = 0; p < g.nGroups; p++) {
        std::vector<V3> dep((size_t)g.maxDeps + 1, V3{0, 0, 0});
        std::vector<int> hit((size_t)g.maxDeps + 1, 0);
        int nDep = 0;
        std::vector<std::vector<V3>> up(4 * 64, std::vector<V3>(15, V3{0, 0, 0}));
        // the group's halo: outside rows, strictly increasing; a meta slot outside it is an error
        const int hb = g.grpHalo[2 * p], hn = g.grpHalo[2 * p + 1];
        for (int h = 1; h < hn; h++) bad += g.haloRows[hb + h] <= g.haloRows[hb + h - 1];
        for (int h = 0; h < hn; h++)
            for (int w = 0; w < 4; w++) {
                const int c = g.runChunks[4 * p + w];
                if (c >= 0) bad += g.haloRows[hb + h] >= t.chunkRow0[c] && g.haloRows[hb + h] < t.chunkRow1[c];
            }
        auto haloRow = [&](int h) {
            if (h < 0 || h >= hn) { bad++; return 0; }
            return g.haloRows[hb + h];
        };
        for (int w = 0; w < 4; w++) {
            const int c = g.runChunks[4 * p + w];
            if (c < 0) continue;
            for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) nDep += t.tPtr[r + 1] - t.tPtr[r];
            const int base = g.chunkUp[2 * c], width = g.chunkUp[2 * c + 1];
            for (int k = 0; k < width; k++)
                for (int lane = 0; lane < 64; lane++) {
                    const size_t q = (size_t)base + (size_t)k * 64 + lane;
                    const int meta = g.upMeta[q];
                    if (meta == -1) continue;
                    const int r = t.chunkRow0[c] + lane;
                    const double *e = vals + g.upSrc[q];
                    if (meta < 0) {
                        const int lc = (meta >> 14) & 0xFF, d = meta & 0x3FFF;
                        const int cc = g.runChunks[4 * p + lc / 64];
                        const int j = t.chunkRow0[cc] + lc % 64;
                        up[w * 64 + lane][k] = prod(e, x + 3 * j, 0);
                        dep[d] = prod(e, x + 3 * r, 1);
                        hit[d]++;
                    } else {
                        up[w * 64 + lane][k] = prod(e, x + 3 * haloRow(meta), 0);
                    }
                }
        }
        const int xb = g.runX[4 * p], xn = g.runX[4 * p + 1];
        for (int q = xb; q < xb + xn; q++) {
            const int K = haloRow(g.xMeta[2 * q]), d = g.xMeta[2 * q + 1];
            dep[d] = prod(vals + g.xSrc[q], x + 3 * K, 1);
            hit[d]++;
        }
        for (int d = 0; d < nDep; d++) bad += hit[d] != 1;
        for (int w = 0; w < 4; w++) {
            const int c = g.runChunks[4 * p + w];
            if (c < 0) continue;
            for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) {
                const int info = g.rowInfo[r], nUp = info & 15, items = (info >> 4) & 31, db = info >> 9;
                V3 v[16];
                v[0] = prod(vals + t.valOff[r], x + 3 * r, 2);
                for (int s = 1; s < 16; s++) {
                    if (s <= nUp) v[s] = up[w * 64 + (r - t.chunkRow0[c])][s - 1];
                    else if (s < items) v[s] = dep[db + s - nUp - 1];
                    else v[s] = {0, 0, 0};
                }
                const V3 y = tree16(v);
                yGrp[3 * r] = y.a; yGrp[3 * r + 1] = y.b; yGrp[3 * r + 2] = y.c;
            }
        }
    }
    return bad;
}
// the group of every chunk (run position), for the tests' spatial checks
void gs_chunk_groups(int *grp) {
    for (int p = 0; p < g_g.nGroups; p++)
        for (int w = 0; w < 4; w++)
            if (g_g.runChunks[4 * p + w] >= 0) grp[g_g.runChunks[4 * p + w]] = p;
}
}
