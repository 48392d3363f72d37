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

// Partially constrained nodes: the reference's layout of the last DOF triple (truncated blocks
// in the block columns, scalar "remained" columns after them; BlockedSymSparseMatrix's
// constructor, core/sparseMatrix.h:186-274) next to the kernels' padded blocks, and the
// value-level gather lists that replace the block-level map.
template <typename ValueOf, typename TetDofs>
static std::string buildPartialMaps(Topology &t, const ValueOf &valueOf, const TetDofs &tetDofs) {
    const int nbc = t.ndofFree / 3, rem = t.ndofFree % 3, r0 = 3 * nbc;
    // remained entries present: some tet holds both DOFs (row >= column)
    bool remHas[3][3] = {};
    for (int e = 0; e < t.nTets; e++) {
        int g[12];
        tetDofs(e, g);
        for (int a = 0; a < 12; a++)
            for (int b = 0; b < 12; b++)
                if (g[a] >= r0 && g[b] >= g[a]) remHas[g[b] - r0][g[a] - r0] = true;
    }
    std::vector<int64_t> refBlk((size_t)t.nBlocks, -1);
    t.refPcol.assign(1, 0);
    t.refOris.assign(1, 0);
    t.refRows.clear();
    int64_t off = 0;
    for (int J = 0; J < nbc; J++) {
        for (int b = t.rowPtr[J]; b < t.rowPtr[J + 1]; b++) {
            const int I = t.colIdx[b];
            refBlk[b] = off;
            t.refRows.push_back(3 * I);
            off += I == J ? 6 : (I < nbc ? 9 : 3 * rem);
        }
        t.refPcol.push_back((int)t.refRows.size());
        t.refOris.push_back((int)off);
    }
    int remPos[3][3];
    for (int c = 0; c < rem; c++) {
        for (int r = c; r < rem; r++) {
            remPos[r][c] = -1;
            if (!remHas[r][c]) continue;
            remPos[r][c] = (int)off++;
            t.refRows.push_back(r0 + r);
        }
        t.refPcol.push_back((int)t.refRows.size());
        t.refOris.push_back((int)off);
    }
    if (off >= ((int64_t)1 << 31)) return "too many values for int32 indexing";
    t.nValuesRef = off;
    t.refPos.assign((size_t)t.nValues, -1);
    for (int J = 0; J < t.nRows; J++)
        for (int b = t.rowPtr[J]; b < t.rowPtr[J + 1]; b++) {
            const int I = t.colIdx[b];
            const int64_t base = t.blkOff[b];
            if (J == nbc) {   // the padded last triple (J == I): its remained entries
                for (int i = 0; i < rem; i++)
                    for (int j = i; j < rem; j++)
                        if (remPos[j][i] >= 0) t.refPos[base + kDiagPacked[i] + (j - i)] = remPos[j][i];
            } else if (I == J) {
                for (int k = 0; k < 6; k++) t.refPos[base + k] = (int)(refBlk[b] + k);
            } else if (I < nbc) {
                for (int k = 0; k < 9; k++) t.refPos[base + k] = (int)(refBlk[b] + k);
            } else {          // truncated block: column-major with rem rows
                for (int sc = 0; sc < 3; sc++)
                    for (int sr = 0; sr < rem; sr++) t.refPos[base + sc * 3 + sr] = (int)(refBlk[b] + sc * rem + sr);
            }
        }
    // value-level gather: each value sums its element entries in tet order, as the reference's
    // scatter does; the entries of a pair (a, b), a <= b, are stored row-major with a's DOF as
    // the row, so local DOFs m <= n give entry 3 (m % 3) + n % 3 of pair slot 10 e + p(a, b)
    std::vector<std::pair<int64_t, int>> vc;
    std::vector<std::pair<int, int>> dc;
    vc.reserve((size_t)78 * t.nTets);
    dc.reserve((size_t)12 * t.nTets);
    for (int e = 0; e < t.nTets; e++) {
        int g[12];
        tetDofs(e, g);
        for (int m = 0; m < 12; m++) {
            if (g[m] < 0) continue;
            dc.emplace_back(g[m], (4 * e + m / 3) * 3 + m % 3);
            for (int n = m; n < 12; n++) {
                if (g[n] < 0) continue;
                const int a = m / 3, bN = n / 3;
                const int p = a * 4 - (a * (a - 1)) / 2 + (bN - a);
                const int64_t v = valueOf(std::max(g[m], g[n]), std::min(g[m], g[n]));
                if (v < 0) return "internal error: element entry outside the pattern";
                vc.emplace_back(v, (10 * e + p) * 9 + 3 * (m % 3) + n % 3);
            }
        }
    }
    if ((int64_t)10 * t.nTets * 9 >= ((int64_t)1 << 31)) return "too many tets for the value-level gather";
    for (int d = t.ndofFree; d < t.ndof; d++) vc.emplace_back(valueOf(d, d), -1);
    std::stable_sort(vc.begin(), vc.end(), [](const std::pair<int64_t, int> &x, const std::pair<int64_t, int> &y) {
        return x.first < y.first;
    });
    std::stable_sort(dc.begin(), dc.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
        return x.first < y.first;
    });
    t.vPtr.assign((size_t)t.nValues + 1, 0);
    t.vSrc.resize(vc.size());
    for (size_t k = 0; k < vc.size(); k++) { t.vPtr[vc[k].first + 1]++; t.vSrc[k] = vc[k].second; }
    for (int64_t v = 0; v < t.nValues; v++) t.vPtr[v + 1] += t.vPtr[v];
    t.dPtr.assign((size_t)t.ndof + 1, 0);
    t.dSrc.resize(dc.size());
    for (size_t k = 0; k < dc.size(); k++) { t.dPtr[dc[k].first + 1]++; t.dSrc[k] = dc[k].second; }
    for (int d = 0; d < t.ndof; d++) t.dPtr[d + 1] += t.dPtr[d];
    for (int d = 0; d < t.ndofFree; d++)
        if (t.dPtr[d + 1] == t.dPtr[d]) {
            char buf[96];
            std::snprintf(buf, sizeof(buf), "free DOF %d belongs to no tet (empty matrix column)", d);
            return buf;
        }
    return "";
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

    // ---- DOF indexing: block row R = free DOFs 3R .. 3R+2 --------------------------
    std::vector<int> nCon(t.nNodes, 0);
    for (int c : t.constrained) nCon[c / 3]++;
    t.partial = false;
    for (int n = 0; n < t.nNodes; n++)
        if (nCon[n] == 1 || nCon[n] == 2) { t.partial = true; break; }
    t.nodeRow.assign(t.nNodes, -1);
    t.rowNode.clear();
    t.nodeDof.clear();
    if (!t.partial) {
        for (int n = 0; n < t.nNodes; n++)
            if (nCon[n] == 0) { t.nodeRow[n] = (int)t.rowNode.size(); t.rowNode.push_back(n); }
        t.nRows = (int)t.rowNode.size();
        t.ndofFree = 3 * t.nRows;
    } else {
        if (t.instNodeStart.size() > 2) return "partially constrained nodes need a single-instance mesh";
        t.nodeDof.assign(3 * (size_t)t.nNodes, -1);
        int d = 0;
        size_t k = 0;
        for (int i = 0; i < 3 * t.nNodes; i++) {
            if (k < t.constrained.size() && t.constrained[k] == i) { k++; continue; }
            if (d % 3 == 0) t.rowNode.push_back(i / 3);
            if (t.nodeRow[i / 3] < 0) t.nodeRow[i / 3] = d / 3;
            t.nodeDof[i] = d++;
        }
        t.ndofFree = d;
        t.nRows = (d + 2) / 3;
    }
    t.ndof = 3 * t.nRows;
    if (t.nRows == 0) return "every node is constrained";
    // free DOFs of a tet's 12 local DOFs (-1: constrained)
    auto tetDofs = [&t](int e, int *g) {
        for (int a = 0; a < 4; a++) {
            const int n = t.tets[4 * (size_t)e + a];
            for (int ax = 0; ax < 3; ax++)
                g[3 * a + ax] = t.partial ? t.nodeDof[3 * (size_t)n + ax] : (t.nodeRow[n] < 0 ? -1 : 3 * t.nodeRow[n] + ax);
        }
    };

    // ---- symbolic pattern: union over tets of free node pairs --------------
    std::vector<uint64_t> keys;
    keys.reserve((size_t)t.nTets * 10);
    for (int e = 0; e < t.nTets; e++) {
        // the tet's block rows (a DOF triple per free node, or up to 12 triples when partial)
        int g[12], rows[12], nr = 0;
        tetDofs(e, g);
        for (int q = 0; q < 12; q++) {
            if (g[q] < 0) continue;
            const int r = g[q] / 3;
            if (std::find(rows, rows + nr, r) == rows + nr) rows[nr++] = r;
        }
        for (int a = 0; a < nr; a++)
            for (int b = a; b < nr; b++) {
                uint64_t J = (uint64_t)std::min(rows[a], rows[b]), I = (uint64_t)std::max(rows[a], rows[b]);
                keys.push_back(J * (uint64_t)t.nRows + I);
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
    // value index of the lower-triangle entry (r, c), r >= c, of the kernels' layout
    auto valueOf = [&t](int r, int c) -> int64_t {
        const int J = c / 3, I = r / 3, sc = c % 3, sr = r % 3;
        if (I == J) return t.valOff[J] + kDiagPacked[sc] + (sr - sc);
        const int b = findBlock(t, J, I);
        return b < 0 ? -1 : (int64_t)t.blkOff[b] + sc * 3 + sr;
    };
    if (t.partial) {
        std::string err = buildPartialMaps(t, valueOf, tetDofs);
        if (!err.empty()) return err;
    } else {
        t.refPcol.clear(); t.refOris.clear(); t.refRows.clear(); t.refPos.clear();
        t.vPtr.clear(); t.vSrc.clear(); t.dPtr.clear(); t.dSrc.clear();
        t.nValuesRef = t.nValues;
    }
    t.elemMap78.assign((size_t)78 * t.nTets, -1);
    t.pairBlk.assign((size_t)10 * t.nTets, -1);
    t.pairT.assign((size_t)10 * t.nTets, 0);
    for (int e = 0; e < t.nTets; e++) {
        int gi[12];
        tetDofs(e, gi);
        int *local = &t.elemMap78[(size_t)78 * e];
        for (int subCol = 0; subCol < 12; subCol++) {
            for (int subRow = subCol; subRow < 12; subRow++) {
                int gc = gi[subCol], gr = gi[subRow];
                int li = localMatrixIndex(subRow, subCol);
                if (gc < 0 || gr < 0) { local[li] = -1; continue; }
                const int64_t v = valueOf(std::max(gc, gr), std::min(gc, gr));
                local[li] = t.partial ? t.refPos[v] : (int)v;
            }
        }
        if (t.partial) continue;   // the value-level gathers replace the block map
        const int *tn = &t.tets[4 * e];
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
    t.rowCStart.assign(t.nRows + 1, 0);
    if (t.partial) {
        // every pair and node of every tet has its own slot, in tet order; the value-level
        // lists read them
        t.nPairContrib = (int64_t)10 * t.nTets;
        t.nNodeContrib = (int64_t)4 * t.nTets;
        t.pairSlot.resize((size_t)10 * t.nTets);
        std::iota(t.pairSlot.begin(), t.pairSlot.end(), 0);
        t.nodeSlot.resize((size_t)4 * t.nTets);
        std::iota(t.nodeSlot.begin(), t.nodeSlot.end(), 0);
    } else {
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
