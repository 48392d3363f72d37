// Host-side set-up (runs once per mesh; not on the per-step path).
//
// The canonical layout reproduced here is the reference's
// BlockedSymSparseMatrix<3,3> (core/sparseMatrix.h:186-274): block column J
// holds its packed 6-value diagonal block followed by the 3x3 blocks of the
// rows below it in ascending order, column-major inside each block.  Read by
// block rows of the upper triangle, that is an upper BSR with row-major 3x3
// blocks, which is how the kernels address it.  The 78-entry element map is
// the one getMatrixIndicesPerElement builds (material/fullOrderNonlinear.h:73-110).
#include "od_setup.h"

#include <algorithm>
#include <cstdio>
#include <numeric>

namespace od {

static const int kDiagPacked[3] = {0, 3, 5};   // packed diagonal start of sub-column c

int globalIndex(const Topology &t, int node, int axis) {
    int index = 3 * node + axis;
    if (t.constrained.empty()) return index;
    int offset = int(std::lower_bound(t.constrained.begin(), t.constrained.end(), index) - t.constrained.begin());
    if (offset < (int)t.constrained.size() && t.constrained[offset] == index) return -1;
    return index - offset;
}

// element.h:109-113: packed index of local pair (a, b), a <= b, among 12 DOFs
static inline int localMatrixIndex(int a, int b) {
    if (a > b) std::swap(a, b);
    return a * 12 - (((a - 1) * a) / 2) + b - a;
}

static int findBlock(const Topology &t, int J, int I) {
    const int *b = t.colIdx.data() + t.rowPtr[J];
    const int *e = t.colIdx.data() + t.rowPtr[J + 1];
    const int *p = std::lower_bound(b, e, I);
    if (p == e || *p != I) return -1;
    return int(p - t.colIdx.data());
}

std::string buildTopology(Topology &t, int chunkRows, int nXcd) {
    char buf[256];
    if (t.nNodes <= 0 || t.nTets <= 0) return "mesh has no nodes or no tets";
    for (size_t i = 0; i < t.tets.size(); i++)
        if (t.tets[i] < 0 || t.tets[i] >= t.nNodes) return "tet references a node outside [0, nNodes)";
    for (size_t i = 0; i + 1 < t.constrained.size(); i++)
        if (t.constrained[i] >= t.constrained[i + 1]) return "constrained DOFs must be sorted and unique";
    if (!t.constrained.empty() && (t.constrained.front() < 0 || t.constrained.back() >= 3 * t.nNodes))
        return "constrained DOF out of range";

    // ---- node -> block row; whole-node constraints only -------------------
    std::vector<int> nCon(t.nNodes, 0);
    for (int c : t.constrained) nCon[c / 3]++;
    t.nodeRow.assign(t.nNodes, -1);
    t.rowNode.clear();
    for (int n = 0; n < t.nNodes; n++) {
        if (nCon[n] != 0 && nCon[n] != 3) {
            std::snprintf(buf, sizeof(buf), "node %d is partially constrained; the device layout needs whole-node constraints", n);
            return buf;
        }
        if (nCon[n] == 0) { t.nodeRow[n] = (int)t.rowNode.size(); t.rowNode.push_back(n); }
    }
    t.nRows = (int)t.rowNode.size();
    t.ndof = 3 * t.nRows;
    if (t.nRows == 0) return "every node is constrained";

    // ---- symbolic pattern: union over tets of free node pairs --------------
    std::vector<uint64_t> keys;
    keys.reserve((size_t)t.nTets * 10);
    for (int e = 0; e < t.nTets; e++) {
        const int *tn = &t.tets[4 * e];
        for (int a = 0; a < 4; a++) {
            int ra = t.nodeRow[tn[a]];
            if (ra < 0) continue;
            for (int b = a; b < 4; b++) {
                int rb = t.nodeRow[tn[b]];
                if (rb < 0) continue;
                uint64_t J = (uint64_t)std::min(ra, rb), I = (uint64_t)std::max(ra, rb);
                keys.push_back(J * (uint64_t)t.nRows + I);
            }
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    t.nBlocks = (int64_t)keys.size();
    if (t.nBlocks >= (int64_t)1 << 31) return "too many blocks for int32 indexing";
    t.rowPtr.assign(t.nRows + 1, 0);
    t.colIdx.resize(keys.size());
    for (size_t k = 0; k < keys.size(); k++) {
        int J = int(keys[k] / (uint64_t)t.nRows);
        t.colIdx[k] = int(keys[k] % (uint64_t)t.nRows);
        t.rowPtr[J + 1]++;
    }
    for (int J = 0; J < t.nRows; J++) t.rowPtr[J + 1] += t.rowPtr[J];
    t.valOff.assign(t.nRows + 1, 0);
    t.blkOff.resize(keys.size());
    int64_t nv = 0;
    for (int J = 0; J < t.nRows; J++) {
        int b0 = t.rowPtr[J], b1 = t.rowPtr[J + 1];
        if (b0 == b1 || t.colIdx[b0] != J) {
            std::snprintf(buf, sizeof(buf), "free node %d belongs to no tet (empty matrix column)", t.rowNode[J]);
            return buf;
        }
        t.valOff[J] = (int)nv;
        t.blkOff[b0] = (int)nv;
        nv += 6;
        for (int b = b0 + 1; b < b1; b++) { t.blkOff[b] = (int)nv; nv += 9; }
        if (nv >= ((int64_t)1 << 31)) return "too many values for int32 indexing";
    }
    t.valOff[t.nRows] = (int)nv;
    t.nValues = nv;

    // ---- transposed index (blocks (K, I), K < I, grouped by I) ----------------
    t.tPtr.assign(t.nRows + 1, 0);
    for (int J = 0; J < t.nRows; J++)
        for (int b = t.rowPtr[J] + 1; b < t.rowPtr[J + 1]; b++) t.tPtr[t.colIdx[b] + 1]++;
    for (int I = 0; I < t.nRows; I++) t.tPtr[I + 1] += t.tPtr[I];
    t.tRow.resize(t.tPtr[t.nRows]);
    t.tBlk.resize(t.tPtr[t.nRows]);
    {
        std::vector<int> cur(t.tPtr.begin(), t.tPtr.end() - 1);
        for (int J = 0; J < t.nRows; J++)
            for (int b = t.rowPtr[J] + 1; b < t.rowPtr[J + 1]; b++) {
                int I = t.colIdx[b];
                t.tRow[cur[I]] = J;
                t.tBlk[cur[I]] = b;
                cur[I]++;
            }
        t.tPos.assign(t.nBlocks, -1);
        for (size_t k = 0; k < t.tBlk.size(); k++) t.tPos[t.tBlk[k]] = (int)k;
    }

    // ---- element maps -----------------------------------------------------------
    t.elemMap78.assign((size_t)78 * t.nTets, -1);
    t.pairBlk.assign((size_t)10 * t.nTets, -1);
    t.pairT.assign((size_t)10 * t.nTets, 0);
    for (int e = 0; e < t.nTets; e++) {
        const int *tn = &t.tets[4 * e];
        int gi[12];
        for (int a = 0; a < 4; a++) {
            int r = t.nodeRow[tn[a]];
            for (int ax = 0; ax < 3; ax++) gi[3 * a + ax] = r < 0 ? -1 : 3 * r + ax;
        }
        int *local = &t.elemMap78[(size_t)78 * e];
        for (int subCol = 0; subCol < 12; subCol++) {
            for (int subRow = subCol; subRow < 12; subRow++) {
                int gc = gi[subCol], gr = gi[subRow];
                int li = localMatrixIndex(subRow, subCol);
                if (gc < 0 || gr < 0) { local[li] = -1; continue; }
                int r = std::max(gc, gr), c = std::min(gc, gr);
                int J = c / 3, I = r / 3, sc = c % 3, sr = r % 3;
                if (I == J) local[li] = t.valOff[J] + kDiagPacked[sc] + (sr - sc);
                else local[li] = t.blkOff[findBlock(t, J, I)] + sc * 3 + sr;
            }
        }
        int p = 0;
        for (int a = 0; a < 4; a++)
            for (int b = a; b < 4; b++, p++) {
                int ra = t.nodeRow[tn[a]], rb = t.nodeRow[tn[b]];
                if (ra < 0 || rb < 0) continue;
                t.pairBlk[(size_t)10 * e + p] = findBlock(t, std::min(ra, rb), std::max(ra, rb));
                t.pairT[(size_t)10 * e + p] = ra > rb ? 1 : 0;
            }
    }

    // ---- deterministic gather slots (ordered by destination, then tet, then pair)
    t.blkCStart.assign(t.nBlocks + 1, 0);
    for (size_t k = 0; k < t.pairBlk.size(); k++)
        if (t.pairBlk[k] >= 0) t.blkCStart[t.pairBlk[k] + 1]++;
    for (int64_t b = 0; b < t.nBlocks; b++) t.blkCStart[b + 1] += t.blkCStart[b];
    t.nPairContrib = t.blkCStart[t.nBlocks];
    t.pairSlot.assign(t.pairBlk.size(), -1);
    {
        std::vector<int> cur(t.blkCStart.begin(), t.blkCStart.end() - 1);
        for (size_t k = 0; k < t.pairBlk.size(); k++)
            if (t.pairBlk[k] >= 0) t.pairSlot[k] = cur[t.pairBlk[k]]++;
    }
    t.rowCStart.assign(t.nRows + 1, 0);
    for (int e = 0; e < t.nTets; e++)
        for (int a = 0; a < 4; a++) {
            int r = t.nodeRow[t.tets[4 * e + a]];
            if (r >= 0) t.rowCStart[r + 1]++;
        }
    for (int r = 0; r < t.nRows; r++) t.rowCStart[r + 1] += t.rowCStart[r];
    t.nNodeContrib = t.rowCStart[t.nRows];
    t.nodeSlot.assign((size_t)4 * t.nTets, -1);
    {
        std::vector<int> cur(t.rowCStart.begin(), t.rowCStart.end() - 1);
        for (int e = 0; e < t.nTets; e++)
            for (int a = 0; a < 4; a++) {
                int r = t.nodeRow[t.tets[4 * e + a]];
                if (r >= 0) t.nodeSlot[(size_t)4 * e + a] = cur[r]++;
            }
    }

    // ---- IC(0) levels: level(I) = 1 + max level(K), K < I adjacent ---------------
    t.level.assign(t.nRows, 0);
    int maxLevel = 0;
    for (int I = 0; I < t.nRows; I++) {
        int l = 0;
        for (int k = t.tPtr[I]; k < t.tPtr[I + 1]; k++) l = std::max(l, t.level[t.tRow[k]] + 1);
        t.level[I] = l;
        maxLevel = std::max(maxLevel, l);
    }
    t.nLevels = maxLevel + 1;
    t.levelPtr.assign(t.nLevels + 1, 0);
    for (int I = 0; I < t.nRows; I++) t.levelPtr[t.level[I] + 1]++;
    for (int l = 0; l < t.nLevels; l++) t.levelPtr[l + 1] += t.levelPtr[l];
    t.levelRows.resize(t.nRows);
    {
        std::vector<int> cur(t.levelPtr.begin(), t.levelPtr.end() - 1);
        for (int I = 0; I < t.nRows; I++) t.levelRows[cur[t.level[I]]++] = I;
    }
    // Contiguous schedule: the rows, in order, cut into ranges with no dependency inside a
    // range (a colour-grouped numbering gives one range per colour).  When that needs no more
    // steps than the dependency levels, the ranges are the schedule: a step's rows are then
    // consecutive, and the sweeps index them without the level row list.  Every row still
    // runs after all rows it depends on, so the results are the same bits.
    {
        std::vector<int> cuts = {0};
        int start = 0;
        for (int I = 0; I < t.nRows; I++)
            for (int k = t.tPtr[I]; k < t.tPtr[I + 1]; k++)
                if (t.tRow[k] >= start) { cuts.push_back(I); start = I; break; }
        cuts.push_back(t.nRows);
        const int nRanges = (int)cuts.size() - 1;
        t.levelsContiguous = nRanges <= t.nLevels;
        if (t.levelsContiguous) {
            t.nLevels = nRanges;
            t.levelPtr = cuts;
            for (int l = 0; l < nRanges; l++)
                for (int I = cuts[l]; I < cuts[l + 1]; I++) { t.level[I] = l; t.levelRows[I] = I; }
        }
    }

    // ---- instances and chunks -----------------------------------------------------
    if (t.instNodeStart.empty()) { t.instNodeStart = {0, t.nNodes}; t.instTetStart = {0, t.nTets}; }
    t.nInst = (int)t.instNodeStart.size() - 1;
    if (t.nInst < 1 || (int)t.instTetStart.size() != t.nInst + 1 || t.instNodeStart.back() != t.nNodes ||
        t.instTetStart.back() != t.nTets || t.instNodeStart[0] != 0 || t.instTetStart[0] != 0)
        return "instance ranges must cover all nodes and tets";
    t.instRowStart.assign(t.nInst + 1, 0);
    t.tetInst.assign(t.nTets, 0);
    t.rowInst.assign(t.nRows, 0);
    for (int i = 0; i < t.nInst; i++) {
        if (t.instNodeStart[i + 1] < t.instNodeStart[i] || t.instTetStart[i + 1] < t.instTetStart[i])
            return "instance ranges must be non-decreasing";
        for (int e = t.instTetStart[i]; e < t.instTetStart[i + 1]; e++) {
            t.tetInst[e] = i;
            for (int a = 0; a < 4; a++) {
                int n = t.tets[4 * e + a];
                if (n < t.instNodeStart[i] || n >= t.instNodeStart[i + 1]) return "tet references a node of another instance";
            }
        }
        int r0 = -1;
        for (int n = t.instNodeStart[i]; n < t.instNodeStart[i + 1]; n++)
            if (t.nodeRow[n] >= 0) { r0 = t.nodeRow[n]; break; }
        t.instRowStart[i] = r0;
    }
    t.instRowStart[t.nInst] = t.nRows;
    for (int i = t.nInst - 1; i >= 0; i--)
        if (t.instRowStart[i] < 0) t.instRowStart[i] = t.instRowStart[i + 1];
    t.instChunkStart.assign(t.nInst + 1, 0);
    t.chunkRow0.clear(); t.chunkRow1.clear(); t.chunkInst.clear();
    for (int i = 0; i < t.nInst; i++) {
        t.instChunkStart[i] = (int)t.chunkRow0.size();
        for (int r = t.instRowStart[i]; r < t.instRowStart[i + 1]; r += chunkRows) {
            t.chunkRow0.push_back(r);
            t.chunkRow1.push_back(std::min(r + chunkRows, t.instRowStart[i + 1]));
            t.chunkInst.push_back(i);
        }
        for (int r = t.instRowStart[i]; r < t.instRowStart[i + 1]; r++) t.rowInst[r] = i;
    }
    t.instChunkStart[t.nInst] = (int)t.chunkRow0.size();

    // ---- XCD-aware launch order --------------------------------------------------
    // Workgroup w is dispatched to XCD (w mod nXcd) and each XCD runs its workgroups in
    // increasing w.  Chunks are sorted by (instance, centroid along the instance's longest
    // axis) and cut into nXcd contiguous runs; the kernels give the j-th workgroup of XCD x
    // the next chunks of run x, so every XCD sweeps one spatial slab and neighbouring rows
    // (which share x entries and transposed blocks) meet in the same L2.
    {
        const int nc = (int)t.chunkRow0.size();
        std::vector<double> key(nc, 0.0);
        std::vector<int> axis(t.nInst, 0);
        for (int i = 0; i < t.nInst; i++) {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int n = t.instNodeStart[i]; n < t.instNodeStart[i + 1]; n++)
                for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], t.X[3 * n + a]); hi[a] = std::max(hi[a], t.X[3 * n + a]); }
            int best = 0;
            for (int a = 1; a < 3; a++) if (hi[a] - lo[a] > hi[best] - lo[best]) best = a;
            axis[i] = best;
        }
        for (int c = 0; c < nc; c++) {
            double sum = 0;
            for (int r = t.chunkRow0[c]; r < t.chunkRow1[c]; r++) sum += t.X[3 * t.rowNode[r] + axis[t.chunkInst[c]]];
            key[c] = sum / std::max(1, t.chunkRow1[c] - t.chunkRow0[c]);
        }
        std::vector<int> sorted(nc);
        std::iota(sorted.begin(), sorted.end(), 0);
        std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) {
            if (t.chunkInst[a] != t.chunkInst[b]) return t.chunkInst[a] < t.chunkInst[b];
            return key[a] < key[b];
        });
        t.chunkSorted = sorted;
        t.perRun = (nc + nXcd - 1) / nXcd;
    }
    return "";
}

}  // namespace od
