// Host set-up of the grouped symmetric product (runs once per mesh).
//
// A group is up to kGroupChunks chunks of one instance, chosen greedily by the number of
// blocks they share: a chunk not yet grouped starts a group and pulls in, one at a time, the
// ungrouped chunk with the most blocks to the group's rows.  On a colour-grouped numbering
// whose chunks are compact tiles (the same tile in every colour), a group is one tile's chunks
// and most of its rows' blocks are internal.
#include <algorithm>
#include <cstdio>
#include <numeric>

#include "od_setup.h"

namespace od {

std::string buildGroupLayout(const Topology &t, int chunkRows, int nXcd, GroupLayout &g) {
    g = GroupLayout();
    const int nc = (int)t.chunkRow0.size();
    if (t.partial) { g.why = "partially constrained nodes"; return ""; }
    if (nc == 0) { g.why = "no chunks"; return ""; }
    if (chunkRows * kGroupChunks > 256) return "group rows exceed the local column encoding";
    for (int r = 0; r < t.nRows; r++) {
        const int items = (t.rowPtr[r + 1] - t.rowPtr[r]) + (t.tPtr[r + 1] - t.tPtr[r]);
        if (items > 16) { g.why = "a row has more than 16 blocks"; return ""; }
    }
    std::vector<int> chunkOf(t.nRows);
    for (int c = 0; c < nc; c++)
        for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) chunkOf[r] = c;

    // ---- chunk adjacency: blocks between two chunks --------------------------------
    std::vector<std::vector<std::pair<int, int>>> adj(nc);   // (chunk, blocks)
    {
        std::vector<int> mark(nc, -1), pos(nc, 0);
        for (int c = 0; c < nc; c++) {
            auto &a = adj[c];
            auto add = [&](int d) {
                if (d == c) return;
                if (mark[d] != c) { mark[d] = c; pos[d] = (int)a.size(); a.emplace_back(d, 0); }
                a[pos[d]].second++;
            };
            for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) {
                for (int b = t.rowPtr[r] + 1; b < t.rowPtr[r + 1]; b++) add(chunkOf[t.colIdx[b]]);
                for (int k = t.tPtr[r]; k < t.tPtr[r + 1]; k++) add(chunkOf[t.tRow[k]]);
            }
        }
    }

    // ---- greedy groups ---------------------------------------------------------------
    std::vector<int> grpOf(nc, -1);
    std::vector<std::vector<int>> groups;
    {
        std::vector<int> score(nc, 0), touched;
        for (int c = 0; c < nc; c++) {
            if (grpOf[c] >= 0) continue;
            const int gi = (int)groups.size();
            std::vector<int> mem = {c};
            grpOf[c] = gi;
            touched.clear();
            auto absorb = [&](int m) {
                for (const auto &e : adj[m])
                    if (grpOf[e.first] < 0) {
                        if (score[e.first] == 0) touched.push_back(e.first);
                        score[e.first] += e.second;
                    }
            };
            absorb(c);
            while ((int)mem.size() < kGroupChunks) {
                int best = -1;
                for (int d : touched)
                    if (grpOf[d] < 0 && score[d] > 0 && (best < 0 || score[d] > score[best] || (score[d] == score[best] && d < best)))
                        best = d;
                if (best < 0) break;
                grpOf[best] = gi;
                mem.push_back(best);
                absorb(best);
            }
            for (int d : touched) score[d] = 0;
            std::sort(mem.begin(), mem.end());
            groups.push_back(mem);
        }
    }
    const int ng = (int)groups.size();

    // ---- per-row slots, deposits and the upper streams --------------------------------
    std::vector<int> localOf(t.nRows, 0), depBase(t.nRows, 0);
    g.rowInfo.assign(t.nRows, 0);
    std::vector<int> grpDeps(ng, 0);
    for (int gi = 0; gi < ng; gi++) {
        int dep = 0;
        for (size_t w = 0; w < groups[gi].size(); w++) {
            const int c = groups[gi][w];
            for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) {
                localOf[r] = (int)w * chunkRows + (r - t.chunkRow0[c]);
                const int nUp = t.rowPtr[r + 1] - t.rowPtr[r] - 1, nLo = t.tPtr[r + 1] - t.tPtr[r];
                depBase[r] = dep;
                g.rowInfo[r] = nUp | ((1 + nUp + nLo) << 4) | (dep << 9);
                dep += nLo;
            }
        }
        if (dep >= (1 << 14)) return "too many lower slots in a group";
        grpDeps[gi] = dep;
        g.maxDeps = std::max(g.maxDeps, dep);
    }
    g.chunkUp.assign(2 * (size_t)nc, 0);
    int64_t nUp = 0;
    for (int c = 0; c < nc; c++) {
        int w = 0;
        for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) w = std::max(w, t.rowPtr[r + 1] - t.rowPtr[r] - 1);
        if (nUp + (int64_t)w * chunkRows >= INT32_MAX) return "too many upper-stream positions";
        g.chunkUp[2 * (size_t)c] = (int)nUp;
        g.chunkUp[2 * (size_t)c + 1] = w;
        nUp += (int64_t)w * chunkRows;
    }
    g.nUp = nUp;
    g.upMeta.assign((size_t)std::max<int64_t>(1, nUp), -1);
    g.upSrc.assign((size_t)std::max<int64_t>(1, nUp), -1);
    for (int c = 0; c < nc; c++) {
        const int gi = grpOf[c];
        for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) {
            for (int k = 0; k < t.rowPtr[r + 1] - t.rowPtr[r] - 1; k++) {
                const int b = t.rowPtr[r] + 1 + k, j = t.colIdx[b];
                const size_t p = (size_t)g.chunkUp[2 * (size_t)c] + (size_t)k * chunkRows + (r - t.chunkRow0[c]);
                g.upSrc[p] = t.blkOff[b];
                if (grpOf[chunkOf[j]] == gi) {
                    const int d = depBase[j] + (t.tPos[b] - t.tPtr[j]);
                    g.upMeta[p] = (int)(0x80000000u | ((unsigned)localOf[j] << 14) | (unsigned)d);
                } else {
                    g.upMeta[p] = j;
                }
            }
        }
    }

    // ---- run order: groups by (instance, first spatial position of their chunks) -------
    std::vector<int> sortedPos(nc);
    for (int p = 0; p < nc; p++) sortedPos[t.chunkSorted[p]] = p;
    std::vector<int> order(ng);
    std::iota(order.begin(), order.end(), 0);
    std::vector<int> key(ng);
    for (int gi = 0; gi < ng; gi++) {
        int k = INT32_MAX;
        for (int c : groups[gi]) k = std::min(k, sortedPos[c]);
        key[gi] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int ia = t.chunkInst[groups[a][0]], ib = t.chunkInst[groups[b][0]];
        if (ia != ib) return ia < ib;
        return key[a] < key[b];
    });

    // ---- crossing lists, in run order ---------------------------------------------------
    g.nGroups = ng;
    g.perRun = (ng + nXcd - 1) / nXcd;
    g.runChunks.assign(4 * (size_t)ng, -1);
    g.runX.assign(4 * (size_t)ng, 0);
    int64_t nX = 0;
    for (int p = 0; p < ng; p++) {
        const int gi = order[p];
        const auto &mem = groups[gi];
        for (size_t w = 0; w < mem.size(); w++) g.runChunks[4 * (size_t)p + w] = mem[w];
        const int64_t base = nX;
        for (int c : mem)
            for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++)
                for (int k = t.tPtr[r]; k < t.tPtr[r + 1]; k++) {
                    const int K = t.tRow[k];
                    if (grpOf[chunkOf[K]] == gi) continue;
                    g.xMeta.push_back(K);
                    g.xMeta.push_back(depBase[r] + (k - t.tPtr[r]));
                    g.xSrc.push_back(t.blkOff[t.tBlk[k]]);
                    nX++;
                }
        if (nX >= INT32_MAX) return "too many crossing positions";
        g.runX[4 * (size_t)p] = (int)base;
        g.runX[4 * (size_t)p + 1] = (int)(nX - base);
        g.runX[4 * (size_t)p + 2] = t.chunkInst[mem[0]];
    }
    // ---- halos: per group the outside rows whose x it multiplies (the column rows of its
    // crossing upper blocks and the upper owners of its crossing list), in row order; the
    // metas name a row by its slot in the group's halo, whose x the product stages in LDS
    g.grpHalo.assign(2 * (size_t)ng, 0);
    {
        std::vector<int> slotOf(t.nRows, -1), ext;
        for (int p = 0; p < ng; p++) {
            const int gi = order[p];
            ext.clear();
            for (int c : groups[gi])
                for (int64_t q = g.chunkUp[2 * (size_t)c]; q < g.chunkUp[2 * (size_t)c] + (int64_t)g.chunkUp[2 * (size_t)c + 1] * chunkRows; q++)
                    if (g.upMeta[q] >= 0) ext.push_back(g.upMeta[q]);
            const int xb = g.runX[4 * (size_t)p], xn = g.runX[4 * (size_t)p + 1];
            for (int q = xb; q < xb + xn; q++) ext.push_back(g.xMeta[2 * (size_t)q]);
            std::sort(ext.begin(), ext.end());
            ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
            for (size_t h = 0; h < ext.size(); h++) slotOf[ext[h]] = (int)h;
            for (int c : groups[gi])
                for (int64_t q = g.chunkUp[2 * (size_t)c]; q < g.chunkUp[2 * (size_t)c] + (int64_t)g.chunkUp[2 * (size_t)c + 1] * chunkRows; q++)
                    if (g.upMeta[q] >= 0) g.upMeta[q] = slotOf[g.upMeta[q]];
            for (int q = xb; q < xb + xn; q++) g.xMeta[2 * (size_t)q] = slotOf[g.xMeta[2 * (size_t)q]];
            g.grpHalo[2 * (size_t)p] = (int)g.haloRows.size();
            g.grpHalo[2 * (size_t)p + 1] = (int)ext.size();
            g.maxHalo = std::max(g.maxHalo, (int)ext.size());
            g.haloRows.insert(g.haloRows.end(), ext.begin(), ext.end());
            for (int r : ext) slotOf[r] = -1;
        }
        if (g.haloRows.empty()) g.haloRows.assign(1, 0);
    }
    while (nX % chunkRows) { g.xMeta.push_back(0); g.xMeta.push_back(-1); g.xSrc.push_back(-1); nX++; }
    if (nX == 0) { g.xMeta.assign(2, 0); g.xMeta[1] = -1; g.xSrc.assign(1, -1); }
    g.nX = nX;
    g.ok = true;
    return "";
}

}  // namespace od
