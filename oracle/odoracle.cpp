// ORACLE — test infrastructure only (checker, never the thing measured or shipped).
//
// FP64 CPU restatement of OpenDeformer's implicit-step path, single-threaded like the
// reference unless oro_set_threads asks for threads in the per-tet element evaluation (which
// changes no bits: the scatter stays sequential in tet order).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library.  Each function names the reference file:line it follows
// (paths relative to OpenDeformer/ in the reference tree).
//
// Parity pins (how each part is checked against something other than itself)
// live in tests/test_oracle_*.py: brute-force pattern construction, closed-form
// element properties, numpy SVD/eigh/Cholesky/solve, the modal Newmark closed
// form.  Functions no such check covers are marked "parity unpinned" below.
#include "oro_numer.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <climits>
#include <list>
#include <map>
#include <queue>
#include <set>
#include <unordered_map>
#include <vector>

namespace oro {

// ---------------------------------------------------------------------------
// NodeIndexer — core/nodeIndexer.cpp:8-74 (free-DOF compaction by lower_bound)
// ---------------------------------------------------------------------------
struct NodeIndexer {
    std::vector<int> constrainIndices;  // sorted (constrainer.cpp:74-79 outputs a std::set in order)
    int getGlobalIndex(int node, int axis) const {
        int index = 3 * node + axis;
        if (constrainIndices.empty()) return index;
        int offset = int(std::lower_bound(constrainIndices.begin(), constrainIndices.end(), index) - constrainIndices.begin());
        if (offset < (int)constrainIndices.size() && index == constrainIndices[offset]) return -1;
        else if (offset == 0) return index;
        else return index - offset;
    }
    void getElementNodesGlobalIndices(const int *nodes, int count, int *out) const {
        for (int n = 0; n < count; n++)
            for (int a = 0; a < 3; a++) out[n * 3 + a] = getGlobalIndex(nodes[n], a);
    }
    int getMatrixOrder(int nNodes) const { return 3 * nNodes - (int)constrainIndices.size(); }
};

// ---------------------------------------------------------------------------
// BlockedSymSparseMatrixAssembler<3,3> — core/sparseMatrix.h:90-182
// ---------------------------------------------------------------------------
static const int kBL = 3, kBW = 3;
static const int kDiagIdx3[3] = {0, 3, 5};  // diagIndicesGen(3, i), sparseMatrix.h:85-88

struct Assembler {
    int numColumn, numBlockColumn, numRemainedColumn;
    std::vector<std::map<int, std::vector<Scalar>>> blockEntries;
    std::map<int, Scalar> remainedEntries;
    std::vector<int> remainedEntryCount;
    explicit Assembler(int columnCount)
        : numColumn(columnCount), numBlockColumn(columnCount / kBL),
          numRemainedColumn(columnCount - (columnCount / kBL) * kBL), blockEntries(columnCount / kBL),
          remainedEntryCount(std::max(0, columnCount - (columnCount / kBL) * kBL), 0) {}
    // sparseMatrix.h:102-164
    void addEntry(int row, int column, Scalar data) {
        int blockColumn = column / kBL;
        int columnStart = blockColumn * kBL;
        int subColumn = column - columnStart;
        if (blockColumn < numBlockColumn) {
            if (row - columnStart < kBL) {
                int subRow = row - columnStart;
                int index = subRow - subColumn + kDiagIdx3[subColumn];
                auto &m = blockEntries[blockColumn];
                auto found = m.find(columnStart);
                if (found != m.end()) found->second[index] += data;
                else { std::vector<Scalar> v(6, 0.0); v[index] = data; m.emplace(columnStart, std::move(v)); }
            } else {
                int normalBlockStartRow = columnStart + kBL;
                int blockIndex = (row - normalBlockStartRow) / kBW;
                int blockStartRow = normalBlockStartRow + blockIndex * kBW;
                int subRow = row - blockStartRow;
                int width = numColumn - blockStartRow;
                auto &m = blockEntries[blockColumn];
                if (width >= kBW) {
                    auto found = m.find(blockStartRow);
                    if (found != m.end()) found->second[subColumn * kBW + subRow] += data;
                    else { std::vector<Scalar> v(9, 0.0); v[subColumn * kBW + subRow] = data; m.emplace(blockStartRow, std::move(v)); }
                } else {
                    auto found = m.find(blockStartRow);
                    if (found != m.end()) found->second[subColumn * width + subRow] += data;
                    else { std::vector<Scalar> v(kBL * width, 0.0); v[subColumn * width + subRow] = data; m.emplace(blockStartRow, std::move(v)); }
                }
            }
        } else {
            int subRow = row - columnStart;
            int index = subColumn * numRemainedColumn + subRow;
            auto found = remainedEntries.find(index);
            if (found != remainedEntries.end()) found->second += data;
            else { remainedEntryCount[subColumn]++; remainedEntries.emplace(index, data); }
        }
    }
};

// ---------------------------------------------------------------------------
// BlockedSymSparseMatrix<3,3> layout — core/sparseMatrix.h:183-274, 627-635
// ---------------------------------------------------------------------------
struct Pattern {
    int numColumns = 0, numBlockColumn = 0, numRemainedColumn = 0;
    std::vector<int> blockPcol, blockColumnOris, blockRows;
    std::vector<Scalar> initValues;
    int numEntries() const { return blockColumnOris[numBlockColumn + numRemainedColumn]; }

    void build(const Assembler &as) {
        numColumns = as.numColumn; numBlockColumn = as.numBlockColumn; numRemainedColumn = as.numRemainedColumn;
        blockPcol.assign(numBlockColumn + numRemainedColumn + 1, 0);
        blockColumnOris.assign(numBlockColumn + numRemainedColumn + 1, 0);
        const int regularSize = kBL * kBW, diagSize = (kBL + 1) * kBL / 2;
        int dataCount = 0;
        for (int i = 0; i < numBlockColumn; i++) {
            int blockCount = (int)as.blockEntries[i].size();
            blockPcol[i + 1] = blockPcol[i] + blockCount;
            if (blockCount != 0) {
                int blockDataCount = blockCount * regularSize;
                auto start = as.blockEntries[i].cbegin();
                bool hasDiag = start->first == i * kBL;
                if (hasDiag) blockDataCount -= (regularSize - diagSize);
                auto last = --(as.blockEntries[i].cend());
                if ((!hasDiag || start != last) && numColumns - last->first < kBW)
                    blockDataCount -= (last->first + kBW - numColumns) * kBL;
                dataCount += blockDataCount;
            }
            blockColumnOris[i + 1] = dataCount;
        }
        for (int i = 0; i < numRemainedColumn; i++) {
            int entryCount = as.remainedEntryCount[i];
            blockPcol[i + numBlockColumn + 1] = blockPcol[i + numBlockColumn] + entryCount;
            blockColumnOris[i + numBlockColumn + 1] = blockColumnOris[i + numBlockColumn] + entryCount;
        }
        int blockRowCount = blockPcol[numBlockColumn + numRemainedColumn];
        dataCount += blockRowCount - blockPcol[numBlockColumn];
        initValues.assign(dataCount, 0.0);
        blockRows.assign(blockRowCount, 0);
        int index = 0;
        Scalar *valueIter = initValues.data();
        for (int i = 0; i < numBlockColumn; i++) {
            if (blockPcol[i + 1] != blockPcol[i]) {
                auto pos = as.blockEntries[i].cbegin();
                auto last = as.blockEntries[i].cend();
                auto end = last--;
                if (pos->first == i * kBL) {
                    blockRows[index++] = pos->first;
                    std::memcpy(valueIter, pos->second.data(), sizeof(Scalar) * diagSize);
                    valueIter += diagSize;
                    pos++;
                }
                if (pos != end) {
                    while (pos != last) {
                        blockRows[index++] = pos->first;
                        std::memcpy(valueIter, pos->second.data(), sizeof(Scalar) * regularSize);
                        valueIter += regularSize;
                        pos++;
                    }
                    if (numColumns - pos->first >= kBW) {
                        blockRows[index++] = pos->first;
                        std::memcpy(valueIter, pos->second.data(), sizeof(Scalar) * regularSize);
                        valueIter += regularSize;
                    } else {
                        blockRows[index++] = pos->first;
                        std::memcpy(valueIter, pos->second.data(), sizeof(Scalar) * (numColumns - pos->first) * kBL);
                        valueIter += (numColumns - pos->first) * kBL;
                    }
                }
            }
        }
        int rowOffset = kBL * numBlockColumn;
        for (auto &e : as.remainedEntries) {
            blockRows[index++] = e.first % numRemainedColumn + rowOffset;
            *valueIter++ = e.second;
        }
    }

    // sparseMatrix.h:528-579
    std::vector<std::unordered_map<int, int>> getIndices() const {
        std::vector<std::unordered_map<int, int>> indices(numColumns);
        const int *colPointer = blockPcol.data();
        const int blockBoundary = numBlockColumn * kBL;
        int index = 0;
        for (int column = 0; column < blockBoundary; column += kBL) {
            int blockIndex = *colPointer++;
            int blockEnd = *colPointer;
            if (blockIndex == blockEnd) continue;
            if (blockRows[blockIndex] == column) {
                for (int i = 0; i < kBL; i++)
                    for (int j = i; j < kBL; j++) indices[column + i].emplace(column + j, index++);
                blockIndex += 1;
            }
            if (blockIndex == blockEnd) continue;
            while (blockIndex < blockEnd - 1) {
                int row = blockRows[blockIndex++];
                for (int i = 0; i < kBL; i++)
                    for (int j = 0; j < kBW; j++) indices[column + i].emplace(row + j, index++);
            }
            int row = blockRows[blockIndex];
            int lastWidth = numColumns - row;
            if (lastWidth >= kBW) lastWidth = kBW;
            for (int i = 0; i < kBL; i++)
                for (int j = 0; j < lastWidth; j++) indices[column + i].emplace(row + j, index++);
        }
        for (int column = blockBoundary; column < numColumns; column++) {
            int iter = *colPointer++;
            int end = *colPointer;
            while (iter < end) indices[column].emplace(blockRows[iter++], index++);
        }
        return indices;
    }

    // sparseMatrix.h:454-526
    std::vector<std::vector<std::pair<int, int>>> getFullIndices() const {
        std::vector<std::vector<std::pair<int, int>>> entryPairs(numColumns);
        const int *colPointer = blockPcol.data();
        const int blockBoundary = numBlockColumn * kBL;
        int index = 0;
        for (int column = 0; column < blockBoundary; column += kBL) {
            int blockIndex = *colPointer++;
            int blockEnd = *colPointer;
            if (blockIndex == blockEnd) continue;
            if (blockRows[blockIndex] == column) {
                for (int i = 0; i < kBL; i++) {
                    entryPairs[column + i].push_back(std::make_pair(column + i, index++));
                    for (int j = i + 1; j < kBL; j++) {
                        entryPairs[column + i].push_back(std::make_pair(column + j, index));
                        entryPairs[column + j].push_back(std::make_pair(column + i, index));
                        index += 1;
                    }
                }
                blockIndex += 1;
            }
            if (blockIndex == blockEnd) continue;
            while (blockIndex < blockEnd - 1) {
                int row = blockRows[blockIndex++];
                for (int i = 0; i < kBL; i++)
                    for (int j = 0; j < kBW; j++) {
                        entryPairs[column + i].push_back(std::make_pair(row + j, index));
                        entryPairs[row + j].push_back(std::make_pair(column + i, index));
                        index += 1;
                    }
            }
            int row = blockRows[blockIndex];
            int lastWidth = numColumns - row;
            if (lastWidth >= kBW) lastWidth = kBW;
            for (int i = 0; i < kBL; i++)
                for (int j = 0; j < lastWidth; j++) {
                    entryPairs[column + i].push_back(std::make_pair(row + j, index));
                    entryPairs[row + j].push_back(std::make_pair(column + i, index));
                    index += 1;
                }
        }
        for (int column = blockBoundary; column < numColumns; column++) {
            int iter = *colPointer++;
            int end = *colPointer;
            if (iter == end) continue;
            int row = blockRows[iter];
            if (row == column) { entryPairs[column].push_back(std::make_pair(column, index++)); iter += 1; }
            while (iter < end) {
                row = blockRows[iter++];
                entryPairs[column].push_back(std::make_pair(row, index));
                entryPairs[row].push_back(std::make_pair(column, index));
                index += 1;
            }
        }
        return entryPairs;
    }

    // sparseMatrix.h:413-426
    void getDiagonal(const Scalar *values, Scalar *diags) const {
        int i = 0;
        const Scalar *vals = values;
        for (; i < numBlockColumn; i++) {
            for (int j = 0; j < kBL; j++) *diags++ = vals[kDiagIdx3[j]];
            vals += blockColumnOris[i + 1] - blockColumnOris[i];
        }
        for (; i < numBlockColumn + numRemainedColumn; i++) {
            *diags++ = *vals;
            vals += blockColumnOris[i + 1] - blockColumnOris[i];
        }
    }
};

// ---------------------------------------------------------------------------
// SpMDV — core/numerMethod.h:134-262 (dest += A src, symmetric, accumulating)
// ---------------------------------------------------------------------------
static void SpMDV(const Pattern &mat, const Scalar *valuesIn, const Scalar *src, Scalar *dest) {
    const int blockColumnCount = mat.numBlockColumn;
    const int columnCount = mat.numColumns;
    const Scalar *values = valuesIn;
    const int *blockPcol = mat.blockPcol.data();
    const int *blockRows = mat.blockRows.data();
    const int regularSize = kBL * kBW, diagSize = (kBL + 1) * kBL / 2;
    Scalar destBlock[kBL], srcBlock[kBL];
    const int blockBoundary = kBL * blockColumnCount;
    for (int startColumn = 0; startColumn < blockBoundary; startColumn += kBL) {
        int index = *blockPcol++;
        int end = *blockPcol;
        if (index == end) continue;
        int row = blockRows[index];
        if (row == startColumn) {
            int diagIndex = 0;
            for (int j = 0; j < kBL; j++) {
                int actualRow = row + j;
                Scalar x = src[actualRow];
                Scalar ret = values[diagIndex++] * x;
                for (int k = j + 1; k < kBL; k++) {
                    Scalar val = values[diagIndex++];
                    dest[row + k] += val * x;
                    ret += val * src[row + k];
                }
                dest[actualRow] += ret;
            }
            index++;
            values += diagSize;
        }
        if (index < end) {
            for (int j = 0; j < kBL; j++) { srcBlock[j] = src[startColumn + j]; destBlock[j] = 0; }
        } else continue;
        while (index < end - 1) {
            row = blockRows[index++];
            for (int j = 0, offset = 0; j < kBL; j++, offset += kBW) {
                Scalar x = srcBlock[j];
                Scalar ret = 0;
                for (int k = 0; k < kBW; k++) {
                    int actualRow = row + k;
                    Scalar val = values[offset + k];
                    dest[actualRow] += val * x;
                    ret += val * src[actualRow];
                }
                destBlock[j] += ret;
            }
            values += regularSize;
        }
        row = blockRows[index];
        int mayDegenWidth = columnCount - row;
        if (mayDegenWidth >= kBW) {
            for (int j = 0, offset = 0; j < kBL; j++, offset += kBW) {
                Scalar x = srcBlock[j];
                Scalar ret = 0;
                for (int k = 0; k < kBW; k++) {
                    int actualRow = row + k;
                    Scalar val = values[offset + k];
                    dest[actualRow] += val * x;
                    ret += val * src[actualRow];
                }
                destBlock[j] += ret;
            }
            values += regularSize;
        } else {
            for (int j = 0, offset = 0; j < kBL; j++, offset += mayDegenWidth) {
                Scalar x = srcBlock[j];
                Scalar ret = 0;
                for (int k = 0; k < mayDegenWidth; k++) {
                    int actualRow = row + k;
                    Scalar val = values[offset + k];
                    dest[actualRow] += val * x;
                    ret += val * src[actualRow];
                }
                destBlock[j] += ret;
            }
            values += kBL * mayDegenWidth;
        }
        for (int j = 0; j < kBL; j++) dest[startColumn + j] += destBlock[j];
    }
    for (int i = blockBoundary; i < columnCount; i++) {
        int index = *blockPcol++;
        int end = *blockPcol;
        if (index == end) continue;
        int row = blockRows[index];
        Scalar y = 0, x = src[i];
        if (row == i) { y += x * (*values++); index++; }
        while (index < end) {
            row = blockRows[index++];
            Scalar val = *values++;
            y += val * src[row];
            dest[row] += val * x;
        }
        dest[i] += y;
    }
}

// ---------------------------------------------------------------------------
// System: mesh + indexer + pattern + per-element map
// ---------------------------------------------------------------------------
struct System {
    int nNodes = 0, nTets = 0;
    std::vector<Vec3> X;          // rest positions (mesh.h:37)
    std::vector<Vec3> disp;       // displacements (mesh.h:39)
    std::vector<int> tets;        // 4 per tet
    NodeIndexer indexer;
    int ndof = 0;
    Pattern pat;
    std::vector<int> elemMap;     // 78 per tet (sparseMatrix.cpp:97-118 storage)
    // corotational plastic material (kind 2): per-tet yield criterion, creep ratio and maximum
    // plastic strain, and the plastic strain state, 9 per tet, zero at the start
    // (corotationalPlastic.cpp:8-21, 45-46)
    std::vector<Scalar> plYield, plCreep, plMax, plStrain;
};

// element.h:109-113
static inline int localMatrixIndex(int a, int b) {
    if (a > b) std::swap(a, b);
    int diagIndex = a * 4 * 3 - (((a - 1) * a) / 2);
    return diagIndex + b - a;
}

// fullOrderNonlinear.h:46-71 + sparseMatrix.h:186-274 + fullOrderNonlinear.h:73-110
static void buildSystem(System &s) {
    s.ndof = s.indexer.getMatrixOrder(s.nNodes);
    Assembler as(s.ndof);
    int gi[12];
    for (int e = 0; e < s.nTets; e++) {
        s.indexer.getElementNodesGlobalIndices(&s.tets[4 * e], 4, gi);
        for (int subCol = 0; subCol < 12; subCol++) {
            int globalCol = gi[subCol];
            if (globalCol >= 0) {
                for (int subRow = subCol; subRow < 12; subRow++) {
                    int globalRow = gi[subRow];
                    if (globalRow >= 0) as.addEntry(std::max(globalRow, globalCol), std::min(globalRow, globalCol), 0);
                }
            }
        }
    }
    s.pat.build(as);
    auto matrixIndices = s.pat.getIndices();
    s.elemMap.assign((size_t)78 * s.nTets, 0);
    for (int e = 0; e < s.nTets; e++) {
        int *local = &s.elemMap[(size_t)78 * e];
        s.indexer.getElementNodesGlobalIndices(&s.tets[4 * e], 4, gi);
        for (int subCol = 0; subCol < 12; subCol++) {
            int globalCol = gi[subCol];
            if (globalCol >= 0) {
                for (int subRow = subCol; subRow < 12; subRow++) {
                    int globalRow = gi[subRow];
                    if (globalRow >= 0) {
                        int r = std::max(globalRow, globalCol), c = std::min(globalRow, globalCol);
                        local[localMatrixIndex(subRow, subCol)] = matrixIndices[c][r];
                    } else local[localMatrixIndex(subRow, subCol)] = -1;
                }
            } else {
                for (int subRow = subCol; subRow < 12; subRow++) local[localMatrixIndex(subRow, subCol)] = -1;
            }
        }
    }
}

// tetelement.cpp:6-12
static inline Scalar getTetVolume(const Vec3 &a, const Vec3 &b, const Vec3 &c, const Vec3 &d) {
    Vec3 ab = b - a, ac = c - a, ad = d - a;
    return std::fabs(ab * (ac % ad)) / 6.f;
}

// tetelement.cpp:18-48
static void getTetShapeFunctionDerivatives(const Vec3 &a, const Vec3 &b, const Vec3 &c, const Vec3 &d,
                                           Scalar *dn0, Scalar *dn1, Scalar *dn2, Scalar *dn3) {
    Vec3 da = a - d, db = b - d, dc = c - d, ca = a - c, cb = b - c;
    Vec3 dNa, dNb, dNc, dNd;
    if (da * (db % dc) > Scalar(0)) { dNa = db % dc; dNb = dc % da; dNc = da % db; dNd = cb % ca; }
    else { dNa = dc % db; dNb = da % dc; dNc = db % da; dNd = ca % cb; }
    for (int i = 0; i < 3; i++) { dn0[i] = dNa[i]; dn1[i] = dNb[i]; dn2[i] = dNc[i]; dn3[i] = dNd[i]; }
}

// corotationalHyperelastic.cpp:8-18 (Lamé D = [lambda + 2 mu, lambda, mu])
static void lameD(Scalar E, Scalar nu, Scalar *D) {
    Scalar lambda = nu * E / ((Scalar(1.0) + nu) * (Scalar(1.0) - Scalar(2.0) * nu));
    Scalar mu = E / (Scalar(2.0) * (Scalar(1.0) + nu));
    D[0] = lambda + Scalar(2.0) * mu; D[1] = lambda; D[2] = mu;
}

// tetelement.cpp:501-544 — K0 (10 row-major 3x3 blocks, a <= b) and Dm^-1
static void corotPrecompute(const Vec3 *v, const Scalar *D, Scalar *K0, Scalar *DmInv) {
    Scalar dr[12];
    getTetShapeFunctionDerivatives(v[0], v[1], v[2], v[3], dr, dr + 3, dr + 6, dr + 9);
    Scalar volume = getTetVolume(v[0], v[1], v[2], v[3]);
    Scalar factor = Scalar(1.0) / (Scalar(36.0) * volume);
    int idx = 0;
    for (int a = 0; a < 4; a++) {
        const Scalar *Ba = dr + 3 * a;
        for (int b = a; b < 4; b++) {
            const Scalar *Bb = dr + 3 * b;
            K0[idx + 0] = (D[0] * Ba[0] * Bb[0] + D[2] * (Ba[1] * Bb[1] + Ba[2] * Bb[2])) * factor;
            K0[idx + 1] = (D[1] * Ba[0] * Bb[1] + D[2] * Ba[1] * Bb[0]) * factor;
            K0[idx + 2] = (D[1] * Ba[0] * Bb[2] + D[2] * Ba[2] * Bb[0]) * factor;
            K0[idx + 3] = (D[1] * Ba[1] * Bb[0] + D[2] * Ba[0] * Bb[1]) * factor;
            K0[idx + 4] = (D[0] * Ba[1] * Bb[1] + D[2] * (Ba[0] * Bb[0] + Ba[2] * Bb[2])) * factor;
            K0[idx + 5] = (D[1] * Ba[1] * Bb[2] + D[2] * Ba[2] * Bb[1]) * factor;
            K0[idx + 6] = (D[1] * Ba[2] * Bb[0] + D[2] * Ba[0] * Bb[2]) * factor;
            K0[idx + 7] = (D[1] * Ba[2] * Bb[1] + D[2] * Ba[1] * Bb[2]) * factor;
            K0[idx + 8] = (D[0] * Ba[2] * Bb[2] + D[2] * (Ba[0] * Bb[0] + Ba[1] * Bb[1])) * factor;
            idx += 9;
        }
    }
    Vec3 ab = v[1] - v[0], ac = v[2] - v[0], ad = v[3] - v[0];
    Ten3 DD(ab.x, ac.x, ad.x, ab.y, ac.y, ad.y, ab.z, ac.z, ad.z);
    Ten3 inv = Inverse(DD);
    std::memcpy(DmInv, &inv.m[0][0], sizeof(Scalar) * 9);
}

// tetelement.cpp:546-573
static void corotRotation(const Vec3 *ori, const Vec3 *cur, const Scalar *DmInv, Scalar threshold, Scalar *R) {
    Scalar g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 1; i < 4; i++) {
        Vec3 ax = cur[i] - cur[0];
        for (int j = 0; j < 3; j++) {
            Scalar d = ax[j];
            for (int k = 0; k < 3; k++) g[j * 3 + k] += d * DmInv[(i - 1) * 3 + k];
        }
    }
    Scalar initVolume = getTetVolume(ori[0], ori[1], ori[2], ori[3]);
    Scalar newVolume = getTetVolume(cur[0], cur[1], cur[2], cur[3]);
    if (newVolume >= threshold * initVolume) polarDecompostion3x3(g, R, nullptr);
    else properQRDecompostion3x3(g, R, nullptr);
}

// tetelement.cpp:583-651 — packed 78 stiffness values and 12 nodal forces
static void corotSubStiffness(const Vec3 *ori, const Vec3 *cur, const Scalar *Q, const Scalar *K0in,
                              Scalar *sub, Scalar *nvw) {
    static const int diagIndices[12] = {0, 12, 23, 33, 42, 50, 57, 63, 68, 72, 75, 77};
    for (int i = 0; i < 12; i++) nvw[i] = 0;
    Scalar QK[9], QKT[9], QKQT[9];
    const Scalar *K = K0in;
    for (int a = 0; a < 4; a++) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                QK[i * 3 + j] = Q[i * 3 + 0] * K[0 * 3 + j] + Q[i * 3 + 1] * K[1 * 3 + j] + Q[i * 3 + 2] * K[2 * 3 + j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                QKQT[i * 3 + j] = QK[i * 3 + 0] * Q[j * 3 + 0] + QK[i * 3 + 1] * Q[j * 3 + 1] + QK[i * 3 + 2] * Q[j * 3 + 2];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3 - i; j++) sub[diagIndices[a * 3 + i] + j] = QKQT[i * 3 + (i + j)];
        Vec3 oa = ori[a], ca = cur[a];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) nvw[a * 3 + i] += QKQT[i * 3 + j] * ca[j] - QK[i * 3 + j] * oa[j];
        K += 9;
        for (int b = a + 1; b < 4; b++) {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    QK[i * 3 + j] = Q[i * 3 + 0] * K[0 * 3 + j] + Q[i * 3 + 1] * K[1 * 3 + j] + Q[i * 3 + 2] * K[2 * 3 + j];
                    QKT[i * 3 + j] = Q[i * 3 + 0] * K[j * 3 + 0] + Q[i * 3 + 1] * K[j * 3 + 1] + Q[i * 3 + 2] * K[j * 3 + 2];
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    QKQT[i * 3 + j] = QK[i * 3 + 0] * Q[j * 3 + 0] + QK[i * 3 + 1] * Q[j * 3 + 1] + QK[i * 3 + 2] * Q[j * 3 + 2];
            Vec3 ob = ori[b], cb = cur[b];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    sub[diagIndices[a * 3 + i] + ((b - a) * 3 - i) + j] = QKQT[i * 3 + j];
                    nvw[a * 3 + i] += QKQT[i * 3 + j] * cb[j] - QK[i * 3 + j] * ob[j];
                    nvw[b * 3 + i] += QKQT[j * 3 + i] * ca[j] - QKT[i * 3 + j] * oa[j];
                }
            K += 9;
        }
    }
}

// ---------------------------------------------------------------------------
// Invertible hyperelastic (neo-Hookean) — material/invertibleHyperelastic.cpp,
// material/invertibleNeoHookean.cpp, element/tetelement.cpp:253-499
// ---------------------------------------------------------------------------
static inline int tensorIndex(int i, int j, int k, int l) { return i * 27 + j * 9 + k * 3 + l; }  // tetelement.h:70

// invertibleHyperelastic.cpp:116-245 (threshold = inversion clamp)
static void modifiedDeformGradient(const Scalar *gradient, Scalar threshold, Scalar *diags, Scalar *leftOrtho, Scalar *rightOrtho) {
    Scalar triMat[6], ev[3];
    Ten3 UT, VT;
    const Scalar epsilon = Scalar(1e-8);
    const int di[3] = {0, 3, 5};
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++)
            triMat[di[i] + j - i] = gradient[j * 3 + 0] * gradient[i * 3 + 0] + gradient[j * 3 + 1] * gradient[i * 3 + 1] +
                                    gradient[j * 3 + 2] * gradient[i * 3 + 2];
    eigenSym3x3(triMat, ev, &UT.m[0][0]);
    if (UT.Determinant() < 0.0) for (int i = 0; i < 3; i++) UT(0, i) = -UT(0, i);
    for (int i = 0; i < 3; i++) ev[i] = ev[i] > Scalar(0) ? std::sqrt(ev[i]) : Scalar(0);
    Ten3 F(gradient);
    VT = UT * F;
    int condition = (ev[0] < epsilon) + ((ev[1] < epsilon) << 1) + ((ev[2] < epsilon) << 2);
    switch (condition) {
    case 0: {
        for (int i = 0; i < 3; i++) {
            Scalar inverse = Scalar(1.0) / ev[i];
            for (int j = 0; j < 3; j++) VT(i, j) *= inverse;
        }
        if (VT.Determinant() < 0) {
            int smallest = ev[0] < ev[1] ? (ev[0] < ev[2] ? 0 : 2) : (ev[1] < ev[2] ? 1 : 2);
            for (int i = 0; i < 3; i++) VT(smallest, i) = -VT(smallest, i);
            ev[smallest] = -ev[smallest];
        }
        break;
    }
    case 1: {
        Scalar i1 = Scalar(1.0) / ev[1], i2 = Scalar(1.0) / ev[2];
        for (int j = 0; j < 3; j++) { VT(1, j) *= i1; VT(2, j) *= i2; }
        Vec3 another = Vec3(VT(1, 0), VT(1, 1), VT(1, 2)) % Vec3(VT(2, 0), VT(2, 1), VT(2, 2));
        VT(0, 0) = another.x; VT(0, 1) = another.y; VT(0, 2) = another.z;
        break;
    }
    case 2: {
        Scalar i0 = Scalar(1.0) / ev[0], i2 = Scalar(1.0) / ev[2];
        for (int j = 0; j < 3; j++) { VT(0, j) *= i0; VT(2, j) *= i2; }
        Vec3 another = Vec3(VT(2, 0), VT(2, 1), VT(2, 2)) % Vec3(VT(0, 0), VT(0, 1), VT(0, 2));
        VT(1, 0) = another.x; VT(1, 1) = another.y; VT(1, 2) = another.z;
        break;
    }
    case 3: {
        Scalar inv = Scalar(1.0) / ev[2];
        for (int i = 0; i < 3; i++) VT(2, i) *= inv;
        Vec3 v1, v2;
        coordinateSystem(Vec3(VT(2, 0), VT(2, 1), VT(2, 2)), v1, v2);
        VT(0, 0) = v1.x; VT(0, 1) = v1.y; VT(0, 2) = v1.z;
        VT(1, 0) = v2.x; VT(1, 1) = v2.y; VT(1, 2) = v2.z;
        break;
    }
    case 4: {
        Scalar i0 = Scalar(1.0) / ev[0], i1 = Scalar(1.0) / ev[1];
        for (int j = 0; j < 3; j++) { VT(0, j) *= i0; VT(1, j) *= i1; }
        Vec3 another = Vec3(VT(0, 0), VT(0, 1), VT(0, 2)) % Vec3(VT(1, 0), VT(1, 1), VT(1, 2));
        VT(2, 0) = another.x; VT(2, 1) = another.y; VT(2, 2) = another.z;
        break;
    }
    case 5: {
        Scalar inv = Scalar(1.0) / ev[1];
        for (int i = 0; i < 3; i++) VT(1, i) *= inv;
        Vec3 v1, v2;
        coordinateSystem(Vec3(VT(1, 0), VT(1, 1), VT(1, 2)), v1, v2);
        VT(2, 0) = v1.x; VT(2, 1) = v1.y; VT(2, 2) = v1.z;
        VT(0, 0) = v2.x; VT(0, 1) = v2.y; VT(0, 2) = v2.z;
        break;
    }
    case 6: {
        Scalar inv = Scalar(1.0) / ev[0];
        for (int i = 0; i < 3; i++) VT(0, i) *= inv;
        Vec3 v1, v2;
        coordinateSystem(Vec3(VT(0, 0), VT(0, 1), VT(0, 2)), v1, v2);
        VT(1, 0) = v1.x; VT(1, 1) = v1.y; VT(1, 2) = v1.z;
        VT(2, 0) = v2.x; VT(2, 1) = v2.y; VT(2, 2) = v2.z;
        break;
    }
    default: {
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) VT(i, j) = 0;
        VT(0, 0) = VT(1, 1) = VT(2, 2) = 1.0;
        break;
    }
    }
    for (int i = 0; i < 3; i++) diags[i] = std::max(ev[i], threshold);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) leftOrtho[i * 3 + j] = UT(j, i);
    std::memcpy(rightOrtho, &VT.m[0][0], sizeof(Scalar) * 9);
}

// invertibleHyperelastic.h:22-30
static void getInvariants(const Scalar *s, Scalar *inv) {
    Scalar s0 = s[0] * s[0], s1 = s[1] * s[1], s2 = s[2] * s[2];
    inv[0] = s0 + s1 + s2;
    inv[1] = s0 * s0 + s1 * s1 + s2 * s2;
    inv[2] = s0 * s1 * s2;
}
// invertibleNeoHookean.cpp:11-20
static void neoEnergyGradient(Scalar lambda, Scalar mu, const Scalar *inv, Scalar *g) {
    g[0] = Scalar(0.5) * mu;
    g[1] = Scalar(0);
    g[2] = (Scalar(-0.5) * mu + Scalar(0.25) * lambda * std::log(inv[2])) / inv[2];
}
static void neoEnergyHessian(Scalar lambda, Scalar mu, const Scalar *inv, Scalar *h) {
    for (int i = 0; i < 6; i++) h[i] = 0;
    h[5] = (Scalar(0.5) * mu + Scalar(0.25) * lambda * (Scalar(1.0) - std::log(inv[2]))) / (inv[2] * inv[2]);
}
// invertibleHyperelastic.cpp:247-265
static void piolaKirchhoff(const Scalar *diags, const Scalar *U, const Scalar *VT, const Scalar *inv, const Scalar *g, Scalar *P) {
    Scalar de[3];
    for (int i = 0; i < 3; i++)
        de[i] = Scalar(2.0) * g[0] * diags[i] + Scalar(4.0) * g[1] * diags[i] * diags[i] * diags[i] +
                Scalar(2.0) * g[2] * inv[2] / diags[i];
    Scalar rm[9];
    std::memcpy(rm, VT, sizeof(Scalar) * 9);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) rm[i * 3 + j] *= de[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            P[i * 3 + j] = U[i * 3 + 0] * rm[0 * 3 + j] + U[i * 3 + 1] * rm[1 * 3 + j] + U[i * 3 + 2] * rm[2 * 3 + j];
}
// tetelement.cpp:446-473
static void forceSemidefinite3x3(Scalar mat[6]) {
    Ten3 QT;
    Scalar ev[3];
    eigenSym3x3(mat, ev, &QT.m[0][0]);
    bool negative = false;
    for (int i = 0; i < 3; i++) if (ev[i] < 0) { negative = true; ev[i] = 0.0; }
    if (negative) {
        const Ten3 Q = Transpose(QT);
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) QT(i, j) *= ev[i];
        const Ten3 A = Q * QT;
        mat[0] = A(0, 0); mat[1] = A(0, 1); mat[2] = A(0, 2);
        mat[3] = A(1, 1); mat[4] = A(1, 2); mat[5] = A(2, 2);
    }
}
// tetelement.cpp:475-499
static void forceSemidefinite2x2(Scalar &diag, Scalar &offdiag) {
    Scalar e0 = diag + offdiag, e1 = diag - offdiag;
    int c = (e0 < 0) + ((e1 < 0) << 1);
    switch (c) {
    case 0: break;
    case 1: e1 *= Scalar(0.5); diag = e1; offdiag = -e1; break;
    case 2: e0 *= Scalar(0.5); diag = e0; offdiag = e0; break;
    default: diag = 0.0; offdiag = 0.0; break;
    }
}
// tetelement.cpp:357-444
static void getdPdF(const Scalar *diag, const Scalar *g, const Scalar *h, Scalar dPdF[81]) {
    const Scalar ds[3] = {diag[0] * diag[0], diag[1] * diag[1], diag[2] * diag[2]};
    const Scalar invariant = ds[0] * ds[1] * ds[2];
    const Scalar dc[3] = {diag[0] * diag[1], diag[0] * diag[2], diag[1] * diag[2]};
    Scalar alpha00 = Scalar(2.0) * g[0] + Scalar(8.0) * ds[0] * g[1];
    Scalar alpha01 = Scalar(2.0) * g[0] + Scalar(4.0) * (ds[0] + ds[1]) * g[1];
    Scalar alpha02 = Scalar(2.0) * g[0] + Scalar(4.0) * (ds[0] + ds[2]) * g[1];
    Scalar alpha11 = Scalar(2.0) * g[0] + Scalar(8.0) * ds[1] * g[1];
    Scalar alpha12 = Scalar(2.0) * g[0] + Scalar(4.0) * (ds[1] + ds[2]) * g[1];
    Scalar alpha22 = Scalar(2.0) * g[0] + Scalar(8.0) * ds[2] * g[1];
    Scalar delta00 = invariant * g[2] / ds[0];
    Scalar delta01 = invariant * g[2] / dc[0];
    Scalar delta02 = invariant * g[2] / dc[1];
    Scalar delta11 = invariant * g[2] / ds[1];
    Scalar delta12 = invariant * g[2] / dc[2];
    Scalar delta22 = invariant * g[2] / ds[2];
    Scalar beta00 = Scalar(4.0) * ds[0] * g[1] - Scalar(2.0) * delta00;
    Scalar beta01 = Scalar(4.0) * dc[0] * g[1] - Scalar(2.0) * delta01;
    Scalar beta02 = Scalar(4.0) * dc[1] * g[1] - Scalar(2.0) * delta02;
    Scalar beta11 = Scalar(4.0) * ds[1] * g[1] - Scalar(2.0) * delta11;
    Scalar beta12 = Scalar(4.0) * dc[2] * g[1] - Scalar(2.0) * delta12;
    Scalar beta22 = Scalar(4.0) * ds[2] * g[1] - Scalar(2.0) * delta22;
    const Ten3 hess(h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]);
    const Vec3 vec0(Scalar(2.0) * diag[0], Scalar(4.0) * diag[0] * ds[0], Scalar(2.0) * invariant / diag[0]);
    const Vec3 vec1(Scalar(2.0) * diag[1], Scalar(4.0) * diag[1] * ds[1], Scalar(2.0) * invariant / diag[1]);
    const Vec3 vec2(Scalar(2.0) * diag[2], Scalar(4.0) * diag[2] * ds[2], Scalar(2.0) * invariant / diag[2]);
    Scalar gamma00 = vec0 * (hess * vec0) + Scalar(4.0) * delta00;
    Scalar gamma01 = vec1 * (hess * vec0) + Scalar(4.0) * delta01;
    Scalar gamma02 = vec2 * (hess * vec0) + Scalar(4.0) * delta02;
    Scalar gamma11 = vec1 * (hess * vec1) + Scalar(4.0) * delta11;
    Scalar gamma12 = vec2 * (hess * vec1) + Scalar(4.0) * delta12;
    Scalar gamma22 = vec2 * (hess * vec2) + Scalar(4.0) * delta22;
    Scalar tm[6];
    tm[0] = alpha00 + beta00 + gamma00; tm[1] = gamma01; tm[2] = gamma02;
    tm[3] = alpha11 + beta11 + gamma11; tm[4] = gamma12; tm[5] = alpha22 + beta22 + gamma22;
    forceSemidefinite3x3(tm);
    forceSemidefinite2x2(alpha01, beta01);
    forceSemidefinite2x2(alpha02, beta02);
    forceSemidefinite2x2(alpha12, beta12);
    for (int i = 0; i < 81; i++) dPdF[i] = 0;
    dPdF[tensorIndex(0, 0, 0, 0)] = tm[0]; dPdF[tensorIndex(0, 0, 1, 1)] = tm[1]; dPdF[tensorIndex(0, 0, 2, 2)] = tm[2];
    dPdF[tensorIndex(1, 1, 0, 0)] = tm[1]; dPdF[tensorIndex(1, 1, 1, 1)] = tm[3]; dPdF[tensorIndex(1, 1, 2, 2)] = tm[4];
    dPdF[tensorIndex(2, 2, 0, 0)] = tm[2]; dPdF[tensorIndex(2, 2, 1, 1)] = tm[4]; dPdF[tensorIndex(2, 2, 2, 2)] = tm[5];
    dPdF[tensorIndex(0, 1, 0, 1)] = alpha01; dPdF[tensorIndex(0, 1, 1, 0)] = beta01;
    dPdF[tensorIndex(1, 0, 0, 1)] = beta01; dPdF[tensorIndex(1, 0, 1, 0)] = alpha01;
    dPdF[tensorIndex(0, 2, 0, 2)] = alpha02; dPdF[tensorIndex(0, 2, 2, 0)] = beta02;
    dPdF[tensorIndex(2, 0, 0, 2)] = beta02; dPdF[tensorIndex(2, 0, 2, 0)] = alpha02;
    dPdF[tensorIndex(1, 2, 1, 2)] = alpha12; dPdF[tensorIndex(1, 2, 2, 1)] = beta12;
    dPdF[tensorIndex(2, 1, 1, 2)] = beta12; dPdF[tensorIndex(2, 1, 2, 1)] = alpha12;
}

// Whole neo-Hookean element: tetelement.cpp:253-355 + invertibleHyperelastic.cpp:64-110
static void neoElement(const Vec3 *ori, const Vec3 *cur, Scalar lambda, Scalar mu, Scalar threshold, Scalar *sub, Scalar *nvw) {
    static const int diagIndices[12] = {0, 12, 23, 33, 42, 50, 57, 63, 68, 72, 75, 77};
    Scalar dr[12];
    getTetShapeFunctionDerivatives(ori[0], ori[1], ori[2], ori[3], dr, dr + 3, dr + 6, dr + 9);
    Vec3 ab = ori[1] - ori[0], ac = ori[2] - ori[0], ad = ori[3] - ori[0];
    Ten3 Dm(ab.x, ac.x, ad.x, ab.y, ac.y, ad.y, ab.z, ac.z, ad.z);
    Ten3 DmInv = Inverse(Dm);
    const Scalar *pre = &DmInv.m[0][0];
    // tetelement.cpp:268-280
    Scalar F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 1; i < 4; i++) {
        Vec3 ax = cur[i] - cur[0];
        for (int j = 0; j < 3; j++) {
            Scalar d = ax[j];
            for (int k = 0; k < 3; k++) F[j * 3 + k] += d * pre[(i - 1) * 3 + k];
        }
    }
    Scalar diags[3], U[9], VT[9], invs[3], g[3], h[6], P[9];
    modifiedDeformGradient(F, threshold, diags, U, VT);
    getInvariants(diags, invs);
    neoEnergyGradient(lambda, mu, invs, g);
    neoEnergyHessian(lambda, mu, invs, h);
    piolaKirchhoff(diags, U, VT, invs, g, P);
    // tetelement.cpp:282-342
    Scalar dPdF[81];
    getdPdF(diags, g, h, dPdF);
    Scalar volume = getTetVolume(ori[0], ori[1], ori[2], ori[3]);
    Scalar factor = Scalar(1.0) / (Scalar(36.0) * volume);
    const Ten3 VTm(VT);
    Vec3 dN[4];
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 3; j++) dN[i][j] = dr[i * 3 + j];
        dN[i] = VTm * dN[i];
    }
    const Ten3 Um(U);
    const Ten3 UTm = Transpose(Um);
    Ten3 sm;
    for (int a = 0; a < 4; a++) {
        const Vec3 dNa = dN[a];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Scalar entry = 0.0;
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) entry += dPdF[tensorIndex(i, k, j, l)] * dNa[k] * dNa[l];
                sm(i, j) = entry;
            }
        sm = Um * sm * UTm;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3 - i; j++) sub[diagIndices[a * 3 + i] + j] = sm(i, i + j) * factor;
        for (int b = a + 1; b < 4; b++) {
            const Vec3 dNb = dN[b];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    Scalar entry = 0.0;
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) entry += dPdF[tensorIndex(i, k, j, l)] * dNa[k] * dNb[l];
                    sm(i, j) = entry;
                }
            sm = Um * sm * UTm;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) sub[diagIndices[a * 3 + i] + ((b - a) * 3 - i) + j] = sm(i, j) * factor;
        }
    }
    // tetelement.cpp:344-355
    const Scalar f6 = Scalar(1.0) / Scalar(6.0);
    Ten3 Pm(P);
    for (int i = 0; i < 3; i++) {
        Vec3 vw = Pm * Vec3(dr[i * 3 + 0], dr[i * 3 + 1], dr[i * 3 + 2]) * f6;
        nvw[i * 3 + 0] = vw.x; nvw[i * 3 + 1] = vw.y; nvw[i * 3 + 2] = vw.z;
    }
    for (int i = 0; i < 3; i++) nvw[9 + i] = -(nvw[0 + i] + nvw[3 + i] + nvw[6 + i]);
}

// corotationalPlastic.cpp:124-183.  The elastic strain is sym(S) - I - plastic strain with S
// the factored part of F (polar or QR); the stress returned is R (lambda tr(e) I + 2 mu e).
// Then, when the norm of the elastic strain exceeds the yield criterion, the plastic strain
// creeps towards it and is capped at the maximum norm.  Both norms sum strain[i]^2 over i, j
// < 3 with only i indexing the strain (:159, :172): 3 (e00^2 + e01^2 + e02^2), as the reference.
static void plasticElasticStress(Scalar lambda, Scalar mu, Scalar yield, Scalar creep, Scalar maxStrain,
                                 const Scalar *ortho, const Scalar *factored, Scalar *plastic, Scalar *elasticStress) {
    const Scalar mu2 = Scalar(2) * mu;
    Scalar e[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            e[i * 3 + j] = Scalar(0.5) * (factored[i * 3 + j] + factored[j * 3 + i]) - plastic[i * 3 + j];
    e[0] -= Scalar(1); e[4] -= Scalar(1); e[8] -= Scalar(1);
    Scalar st[9];
    for (int i = 0; i < 9; i++) st[i] = mu2 * e[i];
    const Scalar dg = lambda * (e[0] + e[4] + e[8]);
    st[0] += dg; st[4] += dg; st[8] += dg;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            elasticStress[i * 3 + j] = ortho[i * 3 + 0] * st[0 * 3 + j] + ortho[i * 3 + 1] * st[1 * 3 + j] +
                                       ortho[i * 3 + 2] * st[2 * 3 + j];
    Scalar en = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) en += e[i] * e[i];
    en = std::sqrt(en);
    if (en > yield) {
        const Scalar ratio = creep * (en - yield) / en;
        for (int i = 0; i < 9; i++) plastic[i] += ratio * e[i];
        Scalar pn = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) pn += plastic[i] * plastic[i];
        pn = std::sqrt(pn);
        if (pn > maxStrain) {
            const Scalar f = maxStrain / pn;
            for (int i = 0; i < 9; i++) plastic[i] *= f;
        }
    }
}

// corotationalPlastic.cpp:57-114 for one tet: stiffness as the corotational element
// (tetelement.cpp:664-713), nodal virtual works from the elastic stress (tetelement.cpp:715-729).
static void plasticElement(const Vec3 *ori, const Vec3 *cur, const Scalar *D, Scalar threshold, Scalar yield,
                           Scalar creep, Scalar maxStrain, Scalar *plastic, Scalar *sub, Scalar *nvw) {
    Scalar K0[90], DmInv[9], dr[12];
    corotPrecompute(ori, D, K0, DmInv);
    getTetShapeFunctionDerivatives(ori[0], ori[1], ori[2], ori[3], dr, dr + 3, dr + 6, dr + 9);
    Scalar g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 1; i < 4; i++) {
        Vec3 ax = cur[i] - cur[0];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) g[j * 3 + k] += ax[j] * DmInv[(i - 1) * 3 + k];
    }
    const Scalar v0 = getTetVolume(ori[0], ori[1], ori[2], ori[3]), v1 = getTetVolume(cur[0], cur[1], cur[2], cur[3]);
    Scalar R[9], S[9];
    if (v1 >= threshold * v0) polarDecompostion3x3(g, R, S);
    else properQRDecompostion3x3(g, R, S);
    Scalar P[9];
    plasticElasticStress(D[1], D[2], yield, creep, maxStrain, R, S, plastic, P);
    corotSubStiffness(ori, cur, R, K0, sub, nvw);   // the 78 stiffness values (its forces are replaced)
    const Scalar f6 = Scalar(1.0) / Scalar(6.0);
    for (int i = 0; i < 3; i++) {
        for (int r = 0; r < 3; r++)
            nvw[i * 3 + r] = (P[r * 3 + 0] * dr[i * 3 + 0] + P[r * 3 + 1] * dr[i * 3 + 1] + P[r * 3 + 2] * dr[i * 3 + 2]) * f6;
    }
    for (int r = 0; r < 3; r++) nvw[9 + r] = -(nvw[0 + r] + nvw[3 + r] + nvw[6 + r]);
}

// Material description per tet: kind 0 = corotational, 1 = invertible neo-Hookean,
// 2 = corotational plastic (parameters and state in the System).
struct Materials {
    std::vector<int> kind;
    std::vector<Scalar> E, nu;
    Scalar corotThreshold = 0, neoThreshold = 0.1;
};

// Threads of the element evaluation (oro_set_threads; default 1, as the reference, which has
// no threads).  Each tet's element values depend only on that tet, and the scatter below runs
// in tet order on one thread whatever the count, so the sums are the same bits.
static int g_threads = 1;

// One tet's element values (sub: 78 packed entries, nvw: 12 forces); updates the tet's own
// plastic state.
static void elementValues(System &s, const Materials &m, int e, Scalar *sub, Scalar *nvw) {
    Scalar K0[90], DmInv[9], R[9], D[3];
    const int *tn = &s.tets[4 * e];
    Vec3 ori[4], cur[4];
    for (int a = 0; a < 4; a++) { ori[a] = s.X[tn[a]]; cur[a] = ori[a] + s.disp[tn[a]]; }
    if (m.kind[e] == 0) {
        lameD(m.E[e], m.nu[e], D);
        corotPrecompute(ori, D, K0, DmInv);
        corotRotation(ori, cur, DmInv, m.corotThreshold, R);
        corotSubStiffness(ori, cur, R, K0, sub, nvw);
    } else if (m.kind[e] == 2) {
        lameD(m.E[e], m.nu[e], D);
        plasticElement(ori, cur, D, m.corotThreshold, s.plYield[e], s.plCreep[e], s.plMax[e], &s.plStrain[9 * (size_t)e],
                       sub, nvw);
    } else {
        Scalar lambda = m.nu[e] * m.E[e] / ((Scalar(1.0) + m.nu[e]) * (Scalar(1.0) - Scalar(2.0) * m.nu[e]));
        Scalar mu = m.E[e] / (Scalar(2.0) * (Scalar(1.0) + m.nu[e]));
        for (int i = 0; i < 78; i++) sub[i] = 0;
        neoElement(ori, cur, lambda, mu, m.neoThreshold, sub, nvw);
    }
}

// corotationalHyperelastic.cpp:43-102 and invertibleHyperelastic.cpp:34-114:
// element loop, K.addEntry(map index, value), vws[dof] += force.  Accumulates.  The element
// values of a block of tets are evaluated (on g_threads threads), then scattered in tet order.
static void assemble(System &s, const Materials &m, Scalar *K, Scalar *vws) {
    constexpr int kTetBlock = 16384;
    std::vector<Scalar> subs((size_t)kTetBlock * 78), nvws((size_t)kTetBlock * 12);
    int gi[12];
    for (int e0 = 0; e0 < s.nTets; e0 += kTetBlock) {
        const int e1 = std::min(s.nTets, e0 + kTetBlock);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int e = e0; e < e1; e++)
            elementValues(s, m, e, &subs[(size_t)(e - e0) * 78], &nvws[(size_t)(e - e0) * 12]);
        for (int e = e0; e < e1; e++) {
            const Scalar *sub = &subs[(size_t)(e - e0) * 78], *nvw = &nvws[(size_t)(e - e0) * 12];
            s.indexer.getElementNodesGlobalIndices(&s.tets[4 * e], 4, gi);
            int entryIndex = 0;
            const int *local = &s.elemMap[(size_t)78 * e];
            for (int subCol = 0; subCol < 12; subCol++) {
                if (gi[subCol] >= 0) {
                    for (int subRow = subCol; subRow < 12; subRow++) {
                        if (gi[subRow] >= 0) K[local[entryIndex]] += sub[entryIndex];
                        entryIndex += 1;
                    }
                } else entryIndex += 12 - subCol;
            }
            for (int l = 0; l < 12; l++) if (gi[l] >= 0) vws[gi[l]] += nvw[l];
        }
    }
}

// mechMaterial.cpp:69-98 + tetelement.cpp:51-66 (rho per tet).  Accumulates.
static void massMatrix(const System &s, const Scalar *rho, Scalar *M) {
    Scalar subMass[10];
    for (int e = 0; e < s.nTets; e++) {
        const int *tn = &s.tets[4 * e];
        Scalar volume = getTetVolume(s.X[tn[0]], s.X[tn[1]], s.X[tn[2]], s.X[tn[3]]);
        Scalar off = volume * Scalar(0.05), dg = volume * Scalar(0.1);
        subMass[0] = dg; subMass[1] = off; subMass[2] = off; subMass[3] = off; subMass[4] = dg;
        subMass[5] = off; subMass[6] = off; subMass[7] = dg; subMass[8] = off; subMass[9] = dg;
        const int *local = &s.elemMap[(size_t)78 * e];
        int k = 0;
        for (int a = 0; a < 4; a++)
            for (int b = a; b < 4; b++) {
                Scalar entry = rho[e] * subMass[k++];
                for (int off3 = 0; off3 < 3; off3++) {
                    int i = s.indexer.getGlobalIndex(tn[a], off3), j = s.indexer.getGlobalIndex(tn[b], off3);
                    if (i >= 0 && j >= 0) M[local[localMatrixIndex(a * 3 + off3, b * 3 + off3)]] += entry;
                }
            }
    }
}

// forcer.cpp:8-33 (body force per tet) + forcer.cpp:80-96 (densify); accumulates into vws.
static void bodyForce(const System &s, const Scalar *bodyPerTet, Scalar *vws) {
    for (int e = 0; e < s.nTets; e++) {
        const int *tn = &s.tets[4 * e];
        Scalar volume = getTetVolume(s.X[tn[0]], s.X[tn[1]], s.X[tn[2]], s.X[tn[3]]);
        Scalar factor = volume / Scalar(4.0);
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) {
                Scalar r = factor * bodyPerTet[3 * e + k];
                if (r != 0.0) {
                    int g = s.indexer.getGlobalIndex(tn[j], k);
                    if (g >= 0) vws[g] += r;
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Preconditioners — core/preconditioner.h:12-22
// ---------------------------------------------------------------------------
struct Preconditioner {
    virtual void solve(int width, const Scalar *rhs, Scalar *result) const { std::memcpy(result, rhs, width * sizeof(Scalar)); }
    virtual void Preprocess(const Pattern &, const Scalar *) {}
    virtual ~Preconditioner() {}
};

// solver/jacobiPreconditioner.cpp:6-26
struct JacobiPre : Preconditioner {
    std::vector<Scalar> invDiags;
    void Preprocess(const Pattern &p, const Scalar *v) override {
        invDiags.resize(p.numColumns);
        p.getDiagonal(v, invDiags.data());
        for (int i = 0; i < p.numColumns; i++) invDiags[i] = invDiags[i] > Scalar(0) ? Scalar(1.0) / invDiags[i] : Scalar(1.0);
    }
    void solve(int width, const Scalar *rhs, Scalar *result) const override {
        for (int i = 0; i < width; i++) result[i] = invDiags[i] * rhs[i];
    }
};

// FastSparseVector — core/latool.h:427-512 (slot 0 is a sentinel; iteration starts at 1)
struct FastSparseVector {
    std::vector<int> indicator, indices;
    std::vector<Scalar> values;
    explicit FastSparseVector(int n) : indicator(n, 0), indices(1, 0), values(1, 0.0) {}
    void Add(int index, Scalar val) {
        int ind = indicator[index];
        if (ind != 0) values[ind] += val;
        else emplaceBack(index, val);
    }
    void emplaceBack(int index, Scalar val) {
        indicator[index] = (int)indices.size();
        indices.push_back(index);
        values.push_back(val);
    }
    void Clear() {
        for (size_t i = 1; i < indices.size(); i++) indicator[indices[i]] = 0;
        indices.resize(1); values.resize(1);
    }
    size_t size() const { return indices.size(); }
};

// BlockedSymSparseMatrix::getColumn — sparseMatrix.h:297-365 (FastSparseVector variant)
static void getColumn(const Pattern &p, const Scalar *values, int column, FastSparseVector &vec) {
    vec.Clear();
    const int regularSize = kBL * kBW, diagSize = (kBL + 1) * kBL / 2;
    int blockIndex = column / kBL;
    int blockStartColumn = blockIndex * kBL;
    if (blockIndex < p.numBlockColumn) {
        int offset = column - blockStartColumn;
        int index = p.blockPcol[blockIndex], end = p.blockPcol[blockIndex + 1];
        if (index == end) return;
        const Scalar *vals = values + p.blockColumnOris[blockIndex];
        if (p.blockRows[index] == blockStartColumn) {
            int start = kDiagIdx3[offset];
            for (int i = 0; i < kBL - offset; i++)
                if (vals[start + i] != Scalar(0)) vec.emplaceBack(column + i, vals[start + i]);
            index++;
            vals += diagSize;
        }
        if (index == end) return;
        while (index < end - 1) {
            int row = p.blockRows[index++];
            int start = offset * kBW;
            for (int i = 0; i < kBW; i++)
                if (vals[start + i] != Scalar(0)) vec.emplaceBack(row + i, vals[start + i]);
            vals += regularSize;
        }
        int row = p.blockRows[index];
        int w = p.numColumns - row;
        if (w >= kBW) {
            int start = offset * kBW;
            for (int i = 0; i < kBW; i++)
                if (vals[start + i] != Scalar(0)) vec.emplaceBack(row + i, vals[start + i]);
        } else {
            int start = offset * w;
            for (int i = 0; i < w; i++)
                if (vals[start + i] != Scalar(0)) vec.emplaceBack(row + i, vals[start + i]);
        }
    } else {
        int start = column - blockStartColumn + p.numBlockColumn;
        int rowIndexStart = p.blockPcol[start];
        int entryCount = p.blockPcol[start + 1] - rowIndexStart;
        const int *row = p.blockRows.data() + rowIndexStart;
        const Scalar *val = values + p.blockColumnOris[start];
        for (int i = 0; i < entryCount; i++)
            if (val[i] != Scalar(0)) vec.emplaceBack(row[i], val[i]);
    }
}

// Lower-triangular CSC factor apply shared by both IC forms:
// solver/inCholeskyPreconditioner.cpp:132-162 (forward column sweep, backward row-dot sweep)
struct CscCholesky : Preconditioner {
    std::vector<Scalar> values;
    std::vector<int> rows, pcol;
    void solve(int width, const Scalar *rhs, Scalar *result) const override {
        int end = pcol[1];
        Scalar sub = rhs[0] / values[0];
        result[0] = sub;
        for (int j = 1; j < width; j++) result[j] = rhs[j];
        for (int j = 1; j < end; j++) result[rows[j]] -= values[j] * sub;
        for (int column = 1; column < width; column++) {
            int start = pcol[column], e2 = pcol[column + 1];
            Scalar s2 = result[column] / values[start];
            result[column] = s2;
            for (int row = start + 1; row < e2; row++) result[rows[row]] -= values[row] * s2;
        }
        result[width - 1] /= values[pcol[width - 1]];
        for (int row = width - 2; row >= 0; row--) {
            Scalar dot = 0;
            int start = pcol[row], e2 = pcol[row + 1];
            for (int column = start + 1; column < e2; column++) dot += result[rows[column]] * values[column];
            result[row] = (result[row] - dot) / values[start];
        }
    }
};

// solver/inCholeskyPreconditioner.cpp:8-130 — left-looking threshold IC with
// drop-and-compensate (the reference's "IC").  Parity pinned at epsilon = 0
// (complete Cholesky), by the compensation property LL^T - A PSD, and at epsilon > 0
// without compensation by the drop rule recomputed densely from the factor, column by
// column.  Parity unpinned: which entries are dropped when compensation is on (each
// compensation raises the pivot before the column's later entries are tested, so the
// decisions follow the sparse vector's visiting order).
struct ThresholdIC : CscCholesky {
    Scalar epsilon, relaxedScale;
    ThresholdIC(Scalar eps, Scalar relax) : epsilon(eps), relaxedScale(relax) {}
    void processColumn(int columnIndex, const FastSparseVector &vec, Scalar *diags, std::pair<int, int> *list,
                       std::vector<std::pair<int, Scalar>> &factorized) {
        Scalar diag = diags[columnIndex];
        factorized.clear();
        Scalar epsilon2 = epsilon * epsilon;
        // `while (++pairIter != pairEnd)` starting from cbegin(): skips the first real entry (the diagonal)
        for (size_t it = 2; it < vec.size(); it++) {
            int row = vec.indices[it];
            Scalar val = vec.values[it];
            if (val != Scalar(0)) {
                if (val * val < std::fabs(epsilon2 * diag * diags[row])) {
                    Scalar comp = std::fabs(val) * relaxedScale;
                    diag += comp;
                    diags[row] += comp;
                } else factorized.emplace_back(row, val);
            }
        }
        int start = pcol[columnIndex];
        int end = start + 1 + (int)factorized.size();
        pcol.emplace_back(end);
        std::sort(factorized.begin(), factorized.end(),
                  [](const std::pair<int, Scalar> &l, const std::pair<int, Scalar> &r) { return l.first < r.first; });
        if (!factorized.empty()) {
            int row = factorized[0].first;
            list[columnIndex] = list[row];
            list[row] = std::make_pair(columnIndex, start + 1);
        }
        if (diag <= Scalar(0)) diag = epsilon2 * diag * diag;
        Scalar diagRoot = std::sqrt(diag);
        values.emplace_back(diagRoot);
        rows.emplace_back(columnIndex);
        Scalar invDiagRoot = Scalar(1.0) / diagRoot;
        for (auto pr : factorized) {
            int row = pr.first;
            Scalar val = pr.second * invDiagRoot;
            rows.emplace_back(row);
            values.emplace_back(pr.second * invDiagRoot);
            diags[row] -= val * val;
        }
    }
    void Preprocess(const Pattern &p, const Scalar *mat) override {
        values.clear(); rows.clear(); pcol.clear();
        int n = p.numColumns;
        pcol.emplace_back(0);
        std::vector<Scalar> diags(n);
        p.getDiagonal(mat, diags.data());
        std::vector<std::pair<int, int>> list(n, std::make_pair(-1, -1));
        std::vector<std::pair<int, Scalar>> factorized;
        FastSparseVector w(n);
        getColumn(p, mat, 0, w);
        processColumn(0, w, diags.data(), list.data(), factorized);
        for (int i = 1; i < n - 1; i++) {
            getColumn(p, mat, i, w);
            auto pr = list[i];
            while (pr.first >= 0) {
                int j = pr.first;
                int trueIndex = pr.second;
                int nextIndex = trueIndex + 1;
                Scalar upper = values[trueIndex];
                for (int k = nextIndex; k < pcol[j + 1]; k++) w.Add(rows[k], -upper * values[k]);
                pr = list[j];
                if (nextIndex < pcol[j + 1]) {
                    int row = rows[nextIndex];
                    list[j] = list[row];
                    list[row] = std::make_pair(j, nextIndex);
                }
            }
            processColumn(i, w, diags.data(), list.data(), factorized);
        }
        // inCholeskyPreconditioner.cpp:124-126: last column has no pivot guard (reference quirk)
        values.emplace_back(std::sqrt(diags[n - 1]));
        rows.emplace_back(n - 1);
        pcol.emplace_back(pcol[n - 1] + 1);
    }
};

// IC(0): incomplete Cholesky restricted to the stored (structural) pattern of
// tril(A), left-looking by columns in the reference's CSC form; pivot rule
// d <= 0 -> eps^2 d^2 as inCholeskyPreconditioner.cpp:52.  No reference
// counterpart (the reference's IC is threshold-based); pinned by
// (LL^T)_ij = A_ij on the pattern and exactness on tree/block-tridiagonal patterns.
struct IC0 : CscCholesky {
    Scalar pivotEps;
    explicit IC0(Scalar eps) : pivotEps(eps) {}
    std::vector<int> posOf;  // canonical value index of each CSC slot
    void Preprocess(const Pattern &p, const Scalar *mat) override {
        int n = p.numColumns;
        if (pcol.empty()) {
            auto idx = p.getIndices();
            pcol.assign(n + 1, 0);
            for (int c = 0; c < n; c++) pcol[c + 1] = pcol[c] + (int)idx[c].size();
            rows.assign(pcol[n], 0);
            posOf.assign(pcol[n], 0);
            for (int c = 0; c < n; c++) {
                std::vector<std::pair<int, int>> e(idx[c].begin(), idx[c].end());
                std::sort(e.begin(), e.end());
                for (size_t k = 0; k < e.size(); k++) { rows[pcol[c] + k] = e[k].first; posOf[pcol[c] + k] = e[k].second; }
            }
        }
        values.assign(pcol[n], 0.0);
        // rowList[j]: (k, slot of L(j,k)) for k < j, ascending k
        std::vector<std::vector<std::pair<int, int>>> rowList(n);
        for (int c = 0; c < n; c++)
            for (int q = pcol[c] + 1; q < pcol[c + 1]; q++) rowList[rows[q]].emplace_back(c, q);
        std::vector<Scalar> w(n, 0.0);
        std::vector<int> mark(n, -1);
        Scalar eps2 = pivotEps * pivotEps;
        for (int j = 0; j < n; j++) {
            for (int q = pcol[j]; q < pcol[j + 1]; q++) { w[rows[q]] = mat[posOf[q]]; mark[rows[q]] = j; }
            for (auto &kq : rowList[j]) {
                int k = kq.first;
                Scalar ljk = values[kq.second];
                for (int q = kq.second; q < pcol[k + 1]; q++) {
                    int i = rows[q];
                    if (mark[i] == j) w[i] -= values[q] * ljk;
                }
            }
            Scalar d = w[j];
            if (d <= Scalar(0)) d = eps2 * d * d;
            Scalar root = std::sqrt(d);
            values[pcol[j]] = root;
            Scalar inv = Scalar(1.0) / root;
            for (int q = pcol[j] + 1; q < pcol[j + 1]; q++) values[q] = w[rows[q]] * inv;
        }
    }
    // factor scattered into the canonical value layout (L(i,j) at the position of A(i,j))
    void exportCanonical(Scalar *out) const {
        for (size_t q = 0; q < values.size(); q++) out[posOf[q]] = values[q];
    }
};

// SparseVector — core/latool.h:348-425
struct SparseVec {
    std::vector<int> idx;
    std::vector<Scalar> val;
    void emplaceBack(int i, Scalar v) { idx.push_back(i); val.push_back(v); }
    void Clear() { idx.clear(); val.clear(); }
};
// latool.cpp:102-162 — FastSparseVector * SparseVector dot
static Scalar dotFS(const FastSparseVector &f, const SparseVec &s) {
    Scalar r = 0;
    for (size_t i = 0; i < s.idx.size(); i++) {
        int ind = f.indicator[s.idx[i]];
        r += f.values[ind] * s.val[i];
    }
    return r;
}
// numerMethod.h:281-300
static void SpMSV(const Scalar *values, const std::vector<std::vector<std::pair<int, int>>> &full, const SparseVec &src,
                  FastSparseVector &dest) {
    for (size_t k = 0; k < src.idx.size(); k++) {
        int column = src.idx[k];
        Scalar entry = src.val[k];
        for (auto &pr : full[column]) dest.Add(pr.first, values[pr.second] * entry);
    }
}

// solver/inLDLTPreconditioner.cpp:8-184 — SAINV/RIF incomplete LDL^T with two
// drop tolerances (2-argument form; the 3-argument constructor self-initialises
// its members, inLDLTPreconditioner.cpp:9).  Pinned at both eps = 0 (exact LDL^T) and,
// with sainvEpsilon = 0, at ldltEpsilon > 0 (the exact factor with its small entries
// dropped).  Parity unpinned: the inverse-factor drops at sainvEpsilon > 0.
struct SainvLDLT : Preconditioner {
    Scalar sainvEpsilon, ldltEpsilon;
    std::vector<Scalar> values, invDiagonal;
    std::vector<int> rows, pcol;
    std::vector<std::vector<std::pair<int, int>>> full;
    SainvLDLT(Scalar s, Scalar l) : sainvEpsilon(s), ldltEpsilon(l) {}
    struct Node { int row, col; Scalar value; Node *prev, *next; };
    void Preprocess(const Pattern &p, const Scalar *mat) override {
        if (full.empty()) full = p.getFullIndices();
        values.clear(); rows.clear(); pcol.clear();
        int n = p.numColumns;
        invDiagonal.assign(n, 0.0);
        std::vector<Node> heads(n);
        std::vector<std::vector<Node *>> invRowNodes(n);
        std::vector<Node *> pool;
        auto alloc = [&]() { Node *nd = new Node(); pool.push_back(nd); return nd; };
        FastSparseVector vec(n), lowerColumn(n);
        SparseVec invColumn;
        for (int i = 0; i < n; i++) {
            Node *nd = alloc();
            nd->row = i; nd->col = i; nd->value = 1.0;
            heads[i].prev = nd; heads[i].next = nd;
            nd->prev = &heads[i]; nd->next = &heads[i];
            invRowNodes[i].push_back(nd);
        }
        int count = 0;
        pcol.push_back(0);
        for (int i = 0; i < n - 1; i++) {
            for (auto nd : invRowNodes[i]) {
                invColumn.emplaceBack(nd->row, nd->value);
                nd->prev->next = nd->next; nd->next->prev = nd->prev;
            }
            invRowNodes[i].clear();
            SpMSV(mat, full, invColumn, vec);
            Scalar diag = dotFS(vec, invColumn);
            Scalar inv = Scalar(1.0) / diag;
            invDiagonal[i] = inv;
            for (size_t it = 1; it < vec.size(); it++) {
                int index = vec.indices[it];
                Scalar value = vec.values[it];
                if (value != 0) {
                    Node *nd = heads[index].next;
                    Node *end = &heads[index];
                    while (nd != end) { lowerColumn.Add(nd->col, value * nd->value); nd = nd->next; }
                }
            }
            for (size_t it = 1; it < lowerColumn.size(); it++) {
                int j = lowerColumn.indices[it];
                Scalar lower = lowerColumn.values[it] * inv;
                if (std::fabs(lower) > ldltEpsilon) { values.push_back(lower); rows.push_back(j); ++count; }
            }
            pcol.push_back(count);
            vec.Clear();
            for (size_t it = 1; it < lowerColumn.size(); it++) {
                int j = lowerColumn.indices[it];
                Scalar factor = lowerColumn.values[it];
                if (factor != 0) {
                    factor *= inv;
                    for (auto nd : invRowNodes[j]) {
                        vec.emplaceBack(nd->row, nd->value);
                        nd->prev->next = nd->next; nd->next->prev = nd->prev;
                    }
                    invRowNodes[j].clear();
                    for (size_t k = 0; k < invColumn.idx.size(); k++) vec.Add(invColumn.idx[k], -factor * invColumn.val[k]);
                    for (size_t k2 = 1; k2 < vec.size(); k2++) {
                        int index = vec.indices[k2];
                        Scalar val = vec.values[k2];
                        if (std::fabs(val) > sainvEpsilon) {
                            Node *nd = alloc();
                            nd->row = index; nd->col = j; nd->value = val;
                            nd->next = heads[index].next; nd->prev = &heads[index];
                            heads[index].next->prev = nd; heads[index].next = nd;
                            invRowNodes[j].push_back(nd);
                        }
                    }
                    vec.Clear();
                }
            }
            invColumn.Clear();
            lowerColumn.Clear();
        }
        int last = n - 1;
        for (auto nd : invRowNodes[last]) invColumn.emplaceBack(nd->row, nd->value);
        SpMSV(mat, full, invColumn, vec);
        invDiagonal[last] = Scalar(1.0) / dotFS(vec, invColumn);
        vec.Clear();
        invColumn.Clear();
        for (auto nd : pool) delete nd;
    }
    // inLDLTPreconditioner.cpp:151-184
    void solve(int width, const Scalar *rhs, Scalar *result) const override {
        int end = pcol[1];
        Scalar sub = rhs[0];
        result[0] = sub;
        for (int j = 1; j < width; j++) result[j] = rhs[j];
        for (int j = 0; j < end; j++) { int row = rows[j]; result[row] = rhs[row] - values[j] * sub; }
        for (int column = 1; column < width - 1; column++) {
            int start = pcol[column], e2 = pcol[column + 1];
            Scalar s2 = result[column];
            for (int row = start; row < e2; row++) result[rows[row]] -= values[row] * s2;
        }
        for (int i = 0; i < width; i++) result[i] *= invDiagonal[i];
        for (int row = width - 2; row >= 0; row--) {
            Scalar dot = 0;
            int start = pcol[row], e2 = pcol[row + 1];
            for (int column = start; column < e2; column++) dot += result[rows[column]] * values[column];
            result[row] -= dot;
        }
    }
};

enum { PRE_NONE = 0, PRE_JACOBI = 1, PRE_ICT = 2, PRE_IC0 = 3, PRE_LDLT = 4 };
static Preconditioner *makePre(int kind, Scalar p1, Scalar p2) {
    switch (kind) {
    case PRE_JACOBI: return new JacobiPre();
    case PRE_ICT: return new ThresholdIC(p1, p2);
    case PRE_IC0: return new IC0(p1);
    case PRE_LDLT: return new SainvLDLT(p1, p2);
    default: return new Preconditioner();
    }
}

// solver/cgSolver.cpp:13-57 — PCG, warm start from result, stop at ||r||^2 <= tol^2 ||b||^2
struct CG {
    Scalar tol; int maxIter; Preconditioner *pre;
    int lastIters = 0; Scalar lastRelRes = 0;
    int solve(const Pattern &p, const Scalar *mat, const Scalar *rhs, Scalar *result) {
        const int width = p.numColumns;
        std::vector<Scalar> memory(3 * (size_t)width, 0.0);
        Scalar *rem = memory.data(), *dir = rem + width, *temp = rem + 2 * width;
        SpMDV(p, mat, result, rem);
        for (int i = 0; i < width; i++) rem[i] = rhs[i] - rem[i];
        pre->Preprocess(p, mat);
        pre->solve(width, rem, dir);
        Scalar delta = Dot(width, rem, dir);
        Scalar bb = Dot(width, rhs, rhs);
        Scalar epsilon = bb * tol * tol;
        Scalar rn2 = Dot(width, rem, rem);
        int iter = 0, done = 0;
        while (iter++ < maxIter && rn2 > epsilon) {
            for (int i = 0; i < width; i++) temp[i] = 0;
            SpMDV(p, mat, dir, temp);
            Scalar alpha = delta / Dot(width, dir, temp);
            for (int i = 0; i < width; i++) { result[i] += alpha * dir[i]; rem[i] -= alpha * temp[i]; }
            pre->solve(width, rem, temp);
            Scalar preDelta = delta;
            delta = Dot(width, rem, temp);
            Scalar beta = delta / preDelta;
            for (int i = 0; i < width; i++) dir[i] = beta * dir[i] + temp[i];
            rn2 = Dot(width, rem, rem);
            done++;
        }
        lastIters = done;
        lastRelRes = bb > 0 ? std::sqrt(rn2 / bb) : std::sqrt(rn2);
        return done;
    }
};

// ---------------------------------------------------------------------------
// Plane collision damping — collision/intersection.cpp:12-143,
// collision/penaltyDampingForcer.cpp:8-61
// ---------------------------------------------------------------------------
struct Plane {
    Vec3 n;        // planeNormal (collisionShapes.h:17-20); below the plane: n.x + c < 0
    Scalar c;      // planeConstant
};

static Vec3 normalize(const Vec3 &v) { return v / v.length(); }   // latool.h:528-530

// intersection.cpp:12-17: point where segment pa-pb crosses the plane, t clamped to [0, 1]
static Vec3 segmentPlane(const Plane &pl, const Vec3 &pa, const Vec3 &pb) {
    Vec3 ab = pb - pa;
    Scalar t = (-pl.c - pl.n * pa) / (pl.n * ab);
    t = t < 0 ? Scalar(0) : (t > 1 ? Scalar(1) : t);               // oder.h:121-125
    return pa + t * ab;
}

// intersection.cpp:20-117: centroid of the part of the tet below the plane and the
// direction vector dir (magnitude six times that part's volume, along the normal).
// Points closer than 1e-6 above the plane count as below.
static void planeTetIntersection(const Plane &pl, const Vec3 *p, Vec3 &centroid, Vec3 &dir) {
    const Scalar eps = Scalar(1e-6);
    Vec3 neg[4], pos[4];
    int nn = 0, np = 0;
    for (int i = 0; i < 4; i++) {
        if (pl.n * p[i] + pl.c < eps) neg[nn++] = p[i];
        else pos[np++] = p[i];
    }
    if (nn == 0) {
        Scalar minDis = std::numeric_limits<Scalar>::max();
        for (int i = 0; i < 4; i++) {
            Scalar dis = pl.n * p[i] + pl.c;
            if (dis < minDis) { minDis = dis; centroid = p[i]; }
        }
        dir = Vec3(0, 0, 0);
    } else if (nn == 1) {
        Vec3 q0 = segmentPlane(pl, neg[0], pos[0]), q1 = segmentPlane(pl, neg[0], pos[1]),
             q2 = segmentPlane(pl, neg[0], pos[2]);
        centroid = (neg[0] + q0 + q1 + q2) * Scalar(0.25);
        Vec3 ab = q0 - neg[0], ac = q1 - neg[0], ad = q2 - neg[0];
        Scalar svol = ab * (ac % ad);
        dir = std::fabs(svol) * normalize((ab % ac) + (ac % ad) + (ad % ab));
        if (svol < 0) dir = -dir;
    } else if (nn == 2) {
        Vec3 q0 = segmentPlane(pl, neg[0], pos[0]), q1 = segmentPlane(pl, neg[0], pos[1]);
        Vec3 q2 = segmentPlane(pl, neg[1], pos[0]), q3 = segmentPlane(pl, neg[1], pos[1]);
        Vec3 ab = neg[1] - neg[0];
        // the wedge below the plane as three tets (n0 n1 q2 q3), (n0 q0 q1 q2), (n0 q1 q2 q3)
        Vec3 c0 = neg[0] + neg[1] + q2 + q3;
        Scalar v0 = std::fabs((neg[1] - neg[0]) * ((q2 - neg[0]) % (q3 - neg[0])));
        Vec3 c1 = neg[0] + q0 + q1 + q2;
        Scalar v1 = std::fabs((q0 - neg[0]) * ((q1 - neg[0]) % (q2 - neg[0])));
        Vec3 c2 = neg[0] + q1 + q2 + q3;
        Scalar v2 = std::fabs((q1 - neg[0]) * ((q2 - neg[0]) % (q3 - neg[0])));
        centroid = (c0 * v0 + c1 * v1 + c2 * v2) / (v0 + v1 + v2) * Scalar(0.25);
        dir = (q0 - neg[0]) % (q1 - neg[0]) + (q3 - neg[1]) % (q2 - neg[1]) + ab % (q2 - neg[0]) +
              (q2 - neg[0]) % (q0 - neg[0]) + (q3 - neg[0]) % ab + (q1 - neg[0]) % (q3 - neg[0]);
        dir = (v0 + v1 + v2) * normalize(dir);
        if ((neg[1] - neg[0]) * ((pos[0] - neg[0]) % (pos[1] - neg[0])) < 0) dir = -dir;
    } else if (nn == 3) {
        Vec3 q0 = segmentPlane(pl, neg[0], pos[0]), q1 = segmentPlane(pl, neg[1], pos[0]),
             q2 = segmentPlane(pl, neg[2], pos[0]);
        Vec3 ab = q0 - pos[0], ac = q1 - pos[0], ad = q2 - pos[0];
        Vec3 c = p[0] + p[1] + p[2] + p[3];
        Scalar sv = (neg[0] - pos[0]) * ((neg[1] - pos[0]) % (neg[2] - pos[0]));
        Scalar v = std::fabs(sv);
        Vec3 c0 = pos[0] + q0 + q1 + q2;
        Scalar v0 = std::fabs(ab * (ac % ad));
        // the tet minus the small tet above the plane at pos[0]
        centroid = (v * c - v0 * c0) / (v - v0) * Scalar(0.25);
        dir = (v - v0) * normalize((ab % ac) + (ac % ad) + (ad % ab));
        if (sv > 0) dir = -dir;
    } else {
        centroid = (p[0] + p[1] + p[2] + p[3]) * Scalar(0.25);
        Scalar vol = std::fabs((p[1] - p[0]) * ((p[2] - p[0]) % (p[3] - p[0])));
        dir = vol * pl.n;
    }
}

// intersection.cpp:119-143: unsigned sub-volume ratios (1/4 each for a tet of 6V <= 1e-6)
static void barycentric(const Vec3 *p, const Vec3 &q, Scalar *coords) {
    const Scalar eps = Scalar(1e-6);
    Vec3 ab = p[1] - p[0], ac = p[2] - p[0], ad = p[3] - p[0], bc = p[2] - p[1], bd = p[3] - p[1];
    Vec3 ap = q - p[0], bp = q - p[1];
    Scalar va = std::fabs(bp * (bc % bd)), vb = std::fabs(ap * (ac % ad)), vc = std::fabs(ap * (ab % ad)),
           vd = std::fabs(ap * (ab % ac));
    Scalar vol = std::fabs(ab * (ac % ad));
    if (vol <= eps) { coords[0] = coords[1] = coords[2] = coords[3] = Scalar(0.25); return; }
    Scalar inv = 1 / vol;
    coords[0] = va * inv; coords[1] = vb * inv; coords[2] = vc * inv; coords[3] = vd * inv;
}

// penaltyDampingForcer.cpp:8-61: adds the damping matrix of one (plane, tet) pair into
// values through the tet's 78-entry map.  vel: velocity per mesh node; positions are the
// rest positions plus the System's displacements.  Nothing is added unless the barycentric
// velocity at the centroid moves into the plane (vel . n̂ < 0).  The stored diagonal-block
// entry (i + j, i) receives t(i, j), not t(i + j, i): penaltyDampingForcer.cpp:43 indexes
// t with the packed offset j, and that is kept as is.
static void planeDampingPair(const System &s, const Plane &pl, int e, Scalar coeff, const Vec3 *vel, Scalar *values) {
    static const int diagIdx[12] = {0, 12, 23, 33, 42, 50, 57, 63, 68, 72, 75, 77};
    const int *tn = &s.tets[4 * (size_t)e];
    const int *local = &s.elemMap[78 * (size_t)e];
    Vec3 p[4];
    for (int i = 0; i < 4; i++) p[i] = s.X[tn[i]] + s.disp[tn[i]];
    Vec3 centroid, dir;
    planeTetIntersection(pl, p, centroid, dir);
    Scalar coords[4];
    barycentric(p, centroid, coords);
    Vec3 nd;
    if (dir.length2() != Scalar(0)) nd = normalize(dir);
    Vec3 v;
    for (int i = 0; i < 4; i++) v = v + coords[i] * vel[tn[i]];
    if (!(v * nd < 0)) return;
    Ten3 t;
    Vec3 cd = coeff * dir;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t(i, j) = cd[i] * nd[j];            // latool.h:554-561
    for (int a = 0; a < 4; a++) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3 - i; j++) {
                int idx = diagIdx[a * 3 + i] + j;
                if (local[idx] >= 0) values[local[idx]] += coords[a] * coords[a] * t(i, j);
            }
        for (int b = a + 1; b < 4; b++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    int idx = diagIdx[a * 3 + i] + ((b - a) * 3 - i) + j;
                    if (local[idx] >= 0) values[local[idx]] += coords[a] * coords[b] * t(i, j);
                }
    }
}

// Node velocities from free-DOF velocities, constrained DOFs zero
// (nonlinearImplicitBackwardEuler.h:131-145).
static std::vector<Vec3> nodeVelocities(const System &s, const Scalar *vFree) {
    std::vector<Vec3> vel(s.nNodes);
    size_t ci = 0; int dof = 0;
    for (int n = 0; n < s.nNodes; n++)
        for (int a = 0; a < 3; a++) {
            if (ci < s.indexer.constrainIndices.size() && 3 * n + a == s.indexer.constrainIndices[ci]) ci++;
            else vel[n][a] = vFree[dof++];
        }
    return vel;
}

// ---------------------------------------------------------------------------
// Integrators
// ---------------------------------------------------------------------------
static void scaleValues(std::vector<Scalar> &v, Scalar a) { for (auto &x : v) x *= a; }                       // sparseMatrix.h:289-293
static void addValues(std::vector<Scalar> &v, Scalar a, const std::vector<Scalar> &o) {                       // sparseMatrix.h:281-287
    for (size_t i = 0; i < v.size(); i++) v[i] += a * o[i];
}

// Writes free-DOF displacement d into node displacements (constrained DOFs keep 0).
static void writeDisplacements(System &s, const Scalar *d) {
    int dof = 0;
    size_t ci = 0;
    for (int n = 0; n < s.nNodes; n++)
        for (int a = 0; a < 3; a++) {
            if (ci < s.indexer.constrainIndices.size() && 3 * n + a == s.indexer.constrainIndices[ci]) ci++;
            else s.disp[n][a] = d[dof++];
        }
}

// Full-order linearised Newmark: one solve per step, the reference's modal
// Newmark update (intergrator/linearNewmark.cpp:27-30, 57-79) lifted to full
// order with the full-order assembly structure of
// intergrator/nonlinearImplicitBackwardEuler.h:82-101.  New code: pinned by
// the modal closed form and discrete energy conservation (beta=1/4, gamma=1/2).
struct Newmark {
    System *s; Materials mat; CG cg; Preconditioner *pre;
    Scalar beta, gamma, h, alphaM, alphaK;
    std::vector<Scalar> d, v, a, fext, impulse, M, K, T, fint, rhs, tmp, dpred, vpred;
    bool initAccel = true; int steps = 0; int lastIters = 0; Scalar lastRelRes = 0; int initIters = 0;
    void initialAcceleration() {
        // a0 = M^-1 (f_ext - f_int(d0) - C v0) ; the full-order form of linearNewmark.cpp:51-55 (a = f at rest)
        int n = s->ndof;
        std::fill(K.begin(), K.end(), 0.0);
        std::fill(fint.begin(), fint.end(), 0.0);
        writeDisplacements(*s, d.data());
        assemble(*s, mat, K.data(), fint.data());
        T = K; scaleValues(T, alphaK); addValues(T, alphaM, M);
        std::fill(rhs.begin(), rhs.end(), 0.0);
        SpMDV(s->pat, T.data(), v.data(), rhs.data());
        for (int i = 0; i < n; i++) rhs[i] = fext[i] - fint[i] - rhs[i];
        JacobiPre jp;
        CG mcg{cg.tol, cg.maxIter, &jp};
        std::fill(a.begin(), a.end(), 0.0);
        initIters = mcg.solve(s->pat, M.data(), rhs.data(), a.data());
    }
    void step() {
        int n = s->ndof;
        if (steps == 0 && initAccel) initialAcceleration();
        const Scalar bh2 = beta * h * h, gh = gamma * h;
        const Scalar mbh2 = h * h * Scalar(0.5) - bh2, mgh = h - gh;   // linearNewmark.cpp:27-30
        for (int i = 0; i < n; i++) {                                  // linearNewmark.cpp:68-69
            dpred[i] = d[i] + h * v[i] + mbh2 * a[i];
            vpred[i] = v[i] + mgh * a[i];
        }
        std::fill(K.begin(), K.end(), 0.0);                            // nonlinearImplicitBackwardEuler.h:83-86
        std::fill(fint.begin(), fint.end(), 0.0);
        writeDisplacements(*s, d.data());
        assemble(*s, mat, K.data(), fint.data());
        T = K; scaleValues(T, alphaK); addValues(T, alphaM, M);       // C = alphaK K + alphaM M
        std::fill(rhs.begin(), rhs.end(), 0.0);
        SpMDV(s->pat, T.data(), vpred.data(), rhs.data());            // C vpred
        for (int i = 0; i < n; i++) tmp[i] = dpred[i] - d[i];
        SpMDV(s->pat, K.data(), tmp.data(), rhs.data());              // + K (dpred - d)
        for (int i = 0; i < n; i++) rhs[i] = fext[i] + impulse[i] - fint[i] - rhs[i];
        scaleValues(T, gh); addValues(T, bh2, K); addValues(T, Scalar(1), M);   // A = M + gh C + bh2 K
        cg.solve(s->pat, T.data(), rhs.data(), a.data());             // warm start from a_n
        lastIters = cg.lastIters; lastRelRes = cg.lastRelRes;
        for (int i = 0; i < n; i++) {                                  // linearNewmark.cpp:76-77
            v[i] = vpred[i] + gh * a[i];
            d[i] = dpred[i] + bh2 * a[i];
        }
        writeDisplacements(*s, d.data());
        std::fill(impulse.begin(), impulse.end(), 0.0);
        steps++;
    }
};

// intergrator/nonlinearImplicitBackwardEuler.h:48-146; D is the external damping matrix
// (addExternalDampingForceMatrix / clearExternalDampingForceMatrix, :122-128)
struct BackwardEuler {
    System *s; Materials mat; CG cg; Preconditioner *pre;
    Scalar h, alphaM, alphaK;
    std::vector<Scalar> v, dv, fint, fext, rhs, K, M, D;
    void step() {
        int n = s->ndof;
        std::fill(K.begin(), K.end(), 0.0);
        std::fill(fint.begin(), fint.end(), 0.0);
        std::fill(rhs.begin(), rhs.end(), 0.0);
        assemble(*s, mat, K.data(), fint.data());
        scaleValues(K, h + alphaK);
        addValues(K, alphaM, M);
        addValues(K, Scalar(1), D);
        SpMDV(s->pat, K.data(), v.data(), rhs.data());
        for (int i = 0; i < n; i++) rhs[i] = h * (fext[i] - fint[i] - rhs[i]);
        scaleValues(K, h);
        addValues(K, Scalar(1), M);
        cg.solve(s->pat, K.data(), rhs.data(), dv.data());
        for (int i = 0; i < n; i++) v[i] += dv[i];
    }
    // nonlinearImplicitBackwardEuler.h:103-116
    void updateMesh() {
        size_t ci = 0; int dof = 0;
        for (int nd = 0; nd < s->nNodes; nd++)
            for (int ax = 0; ax < 3; ax++) {
                if (ci < s->indexer.constrainIndices.size() && 3 * nd + ax == s->indexer.constrainIndices[ci]) ci++;
                else s->disp[nd][ax] += v[dof++] * h;
            }
    }
};

static Materials makeMaterials(int nTets, const int *kind, const Scalar *E, const Scalar *nu, Scalar thrC, Scalar thrN) {
    Materials m;
    m.kind.assign(kind, kind + nTets);
    m.E.assign(E, E + nTets);
    m.nu.assign(nu, nu + nTets);
    m.corotThreshold = thrC; m.neoThreshold = thrN;
    return m;
}

// ---------------------------------------------------------------------------
// GPS node relabelling — core/meshRelabler.cpp:57-530, core/meshRelabeler.h:16-135
//
// The reference leaves four orders to its standard library; the restatement fixes them:
//  (1) edges are visited in the order they are first met (element order, then node pairs
//      j < k), where the reference iterates an unordered_set (meshRelabler.cpp:131-145);
//      each vertex's arc list is built by prepending, as insertArc does (:10-16);
//  (2) std::sort of a level by degree (meshRelabeler.h:85-89) keeps equal degrees in
//      their order (stable);
//  (3) the max-degree priority queue (DegreeComparer, meshRelabeler.h:59-63) pops equal
//      degrees in push order;
//  (4) the sorts of components by size (:334-336) and of elements by label sum (:496)
//      are stable.
// Parity unpinned: the exact labels of the reference binary, which depend on those four
// library orders (pinned instead by hand-traced graphs and the level-structure properties).
// A disconnected mesh has no level structure over all nodes; it is rejected (the
// reference leaves such nodes unnumbered).
// ---------------------------------------------------------------------------
struct GpsGraph {
    std::vector<std::list<int>> adj;   // arc lists (front = last inserted)
    int degree(int v) const { return (int)adj[v].size(); }
};

// RootedLevelStructure::generateLevels (meshRelabler.cpp:57-95): levels[0] = neighbours
// of the root in arc order, then breadth-first in queue order
struct GpsLevels {
    int root = -1;
    std::vector<std::vector<int>> levels;
    int levelSize() const { return (int)levels.size() + 1; }
    int width() const {
        int w = INT_MIN;
        for (auto &l : levels) w = std::max(w, (int)l.size());
        return w;
    }
    void generate(const GpsGraph &g, int r, std::vector<char> &visited) {
        root = r;
        levels.clear();
        visited[r] = 1;
        std::queue<int> q;
        levels.emplace_back();
        for (int u : g.adj[r]) { levels[0].push_back(u); q.push(u); visited[u] = 1; }
        int lv = 0;
        while (!q.empty()) {
            lv++;
            levels.emplace_back();
            int n = (int)q.size();
            for (int i = 0; i < n; i++) {
                int v = q.front(); q.pop();
                for (int u : g.adj[v])
                    if (!visited[u]) { levels[lv].push_back(u); q.push(u); visited[u] = 1; }
            }
        }
        levels.pop_back();
    }
};

// MeshRelabeler::processSigleVert (:148-172)
static void gpsProcess(const GpsGraph &g, int v, std::list<int> &levelNodes, std::queue<int> &working,
                       std::vector<int> &label, int &next) {
    label[v] = next++;
    working.push(v);
    std::vector<int> toBe;                       // priority queue: max degree, ties in push order
    for (int u : g.adj[v]) {
        auto f = std::find(levelNodes.begin(), levelNodes.end(), u);
        if (f != levelNodes.end()) { toBe.push_back(u); levelNodes.erase(f); }
    }
    while (!toBe.empty()) {
        size_t best = 0;
        for (size_t i = 1; i < toBe.size(); i++)
            if (g.degree(toBe[i]) > g.degree(toBe[best])) best = i;
        int u = toBe[best];
        toBe.erase(toBe.begin() + best);
        label[u] = next++;
        working.push(u);
    }
}

// MeshRelabeler::setNewNodeLables (:174-466); returns false for a disconnected graph
static bool gpsLabels(GpsGraph &g, std::vector<int> &out) {
    const int n = (int)g.adj.size();
    std::vector<char> visited(n, 0);
    int start = 0;
    for (int v = 1; v < n; v++)
        if (g.degree(v) < g.degree(start)) start = v;
    GpsLevels startL, endL, mayL;
    startL.generate(g, start, visited);
    {
        int reached = 1;
        for (auto &l : startL.levels) reached += (int)l.size();
        if (reached != n) return false;
    }
    bool found = false;
    while (!found) {
        found = true;
        int endWidth = INT_MAX;
        int last = startL.levelSize() - 1;
        std::vector<int> &lastLevel = startL.levels[last - 1];
        std::stable_sort(lastLevel.begin(), lastLevel.end(),
                         [&](int a, int b) { return g.degree(a) < g.degree(b); });
        for (size_t k = 0; k < lastLevel.size(); k++) {
            std::fill(visited.begin(), visited.end(), 0);
            mayL.generate(g, lastLevel[k], visited);
            if (mayL.levelSize() > startL.levelSize()) {
                found = false;
                std::swap(startL, mayL);
                break;
            }
            int w = mayL.width();
            if (w < endWidth) { std::swap(endL, mayL); endWidth = w; }
        }
    }
    // minimising level width (:224-373)
    const int startSize = startL.levelSize(), endSize = endL.levelSize();
    const int size = std::max(startSize, endSize);
    const int startWidth = startL.width();
    const bool startSmaller = endL.width() > startWidth;
    std::vector<int> sLev(n, 0), eLev(n, INT_MIN);
    sLev[startL.root] = 0;
    for (int i = 1; i < startSize; i++)
        for (int v : startL.levels[i - 1]) sLev[v] = i;
    std::vector<int> deleted;
    std::queue<int> remained;
    std::vector<std::list<int>> newLevels(size);
    auto place = [&](int v, int e) {
        eLev[v] = e;
        if (sLev[v] == e) { deleted.push_back(v); newLevels[e].push_back(v); }
        else remained.push(v);
    };
    int e = endSize - 1;
    place(endL.root, e);
    for (int i = 1; i < endSize; i++) {
        e--;
        for (int v : endL.levels[i - 1]) place(v, e);
    }
    // deleteArc of every deleted vertex from its neighbours' lists (:285-293)
    for (int v : deleted)
        for (int u : g.adj[v]) {
            auto f = std::find(g.adj[u].begin(), g.adj[u].end(), v);
            if (f != g.adj[u].end()) g.adj[u].erase(f);
        }
    std::fill(visited.begin(), visited.end(), 0);
    std::vector<std::vector<int>> comps;
    std::queue<int> working;
    while (!remained.empty()) {
        int v = remained.front();
        if (!visited[v]) {
            comps.emplace_back();
            visited[v] = 1;
            comps.back().push_back(v);
            for (int u : g.adj[v]) working.push(u);
            while (!working.empty()) {
                int a = working.front();
                if (!visited[a]) {
                    visited[a] = 1;
                    comps.back().push_back(a);
                    for (int u : g.adj[a]) working.push(u);
                }
                working.pop();
            }
        }
        remained.pop();
    }
    std::stable_sort(comps.begin(), comps.end(),
                     [](const std::vector<int> &a, const std::vector<int> &b) { return a.size() < b.size(); });
    bool selectSecond = false, interchange = false;
    std::vector<int> c1(size), c2(size);
    for (size_t i = 0; i < comps.size(); i++) {
        std::fill(c1.begin(), c1.end(), 0);
        std::fill(c2.begin(), c2.end(), 0);
        for (int v : comps[i]) { c1[sLev[v]]++; c2[eLev[v]]++; }
        int m1 = *std::max_element(c1.begin(), c1.end()), m2 = *std::max_element(c2.begin(), c2.end());
        if (i == 0 && (m1 > m2 || (m1 == m2 && !startSmaller))) selectSecond = true;
        bool useStart = m1 < m2 ? true : (m1 > m2 ? false : startSmaller);
        for (int v : comps[i]) newLevels[useStart ? sLev[v] : eLev[v]].push_back(v);
    }
    // restore the arcs: insertArc prepends (:375-383)
    for (int v : deleted)
        for (int u : g.adj[v]) g.adj[u].push_front(v);
    for (auto &l : newLevels) {                  // std::list::sort is stable
        std::vector<int> tmp(l.begin(), l.end());
        std::stable_sort(tmp.begin(), tmp.end(), [&](int a, int b) { return g.degree(a) < g.degree(b); });
        l.assign(tmp.begin(), tmp.end());
    }
    // numbering (:389-457)
    int root = startL.root, init = 0, pace = 1;
    if (g.degree(startL.root) > g.degree(endL.root)) { interchange = true; root = endL.root; init = size - 1; pace = -1; }
    if (sLev[root] == eLev[root]) newLevels[init].remove(root);
    else { newLevels[sLev[root]].remove(root); newLevels[eLev[root]].remove(root); }
    std::vector<int> label(n, -1);
    int next = 0;
    std::queue<int> work;
    int level = init;
    gpsProcess(g, root, newLevels[level], work, label, next);
    while (!newLevels[level].empty()) {
        int v = newLevels[level].front();
        newLevels[level].pop_front();
        gpsProcess(g, v, newLevels[level], work, label, next);
    }
    for (level = init + pace; level >= 0 && level < size; level += pace) {
        bool unfound = true;
        int process = -1;
        while (unfound && !work.empty()) {
            int v = work.front();
            for (int u : g.adj[v])
                if (std::find(newLevels[level].begin(), newLevels[level].end(), u) != newLevels[level].end()) {
                    unfound = false;
                    if (process < 0 || g.degree(process) > g.degree(u)) process = u;
                }
            work.pop();
        }
        std::queue<int>().swap(work);
        if (process < 0) {                       // a level no earlier vertex reaches (not in a connected mesh)
            if (newLevels[level].empty()) continue;
            process = newLevels[level].front();
        }
        newLevels[level].remove(process);
        gpsProcess(g, process, newLevels[level], work, label, next);
        while (!newLevels[level].empty()) {
            int v = newLevels[level].front();
            newLevels[level].pop_front();
            gpsProcess(g, v, newLevels[level], work, label, next);
        }
    }
    out.resize(n);
    for (int v = 0; v < n; v++) out[v] = (interchange != selectSecond) ? label[v] : (n - 1) - label[v];
    return true;
}

}  // namespace oro

using namespace oro;

// ---------------------------------------------------------------------------
// C API for the Python tests (ctypes)
// ---------------------------------------------------------------------------
extern "C" {

// constrainer.cpp:57-72 + constrainer.cpp:6-55 + 74-79: scan facets, sorted unique DOFs
int oro_constrain_plane(int nNodes, const double *X, int nFacets, const int *facets, int vertsPerFacet,
                        const double *origin, const double *normal, int type, int *out, int cap) {
    (void)nNodes;
    const Scalar epsilon = Scalar(1e-4);
    std::set<int> c;
    for (int i = 0; i < nFacets; i++)
        for (int j = 0; j < vertsPerFacet; j++) {
            int index = facets[i * vertsPerFacet + j];
            Vec3 vert(X[3 * index], X[3 * index + 1], X[3 * index + 2]);
            Vec3 o(origin[0], origin[1], origin[2]), nrm(normal[0], normal[1], normal[2]);
            if (std::fabs((vert - o) * nrm) < epsilon) {
                switch (type) {  // ConstrainType, constrainer.h:16-24
                case 0: c.insert(index * 3); break;
                case 1: c.insert(index * 3 + 1); break;
                case 2: c.insert(index * 3 + 2); break;
                case 3: c.insert(index * 3); c.insert(index * 3 + 1); break;
                case 4: c.insert(index * 3 + 1); c.insert(index * 3 + 2); break;
                case 5: c.insert(index * 3); c.insert(index * 3 + 2); break;
                default: c.insert(index * 3); c.insert(index * 3 + 1); c.insert(index * 3 + 2); break;
                }
            }
        }
    int k = 0;
    for (int v : c) { if (k < cap) out[k] = v; k++; }
    return k;
}

void *oro_system_create(int nNodes, const double *X, int nTets, const int *tets, int nCon, const int *con) {
    System *s = new System();
    s->nNodes = nNodes; s->nTets = nTets;
    s->X.resize(nNodes); s->disp.assign(nNodes, Vec3());
    for (int i = 0; i < nNodes; i++) s->X[i] = Vec3(X[3 * i], X[3 * i + 1], X[3 * i + 2]);
    s->tets.assign(tets, tets + 4 * (size_t)nTets);
    s->indexer.constrainIndices.assign(con, con + nCon);
    buildSystem(*s);
    // plastic parameters until oro_set_plastic: never yielding (the corotational response)
    s->plYield.assign(nTets, HUGE_VAL); s->plCreep.assign(nTets, 0.0); s->plMax.assign(nTets, HUGE_VAL);
    s->plStrain.assign(9 * (size_t)nTets, 0.0);
    return s;
}
void oro_system_destroy(void *h) { delete (System *)h; }
// threads of the element evaluation (1 = the reference's single thread); returns the count
int oro_set_threads(int n) { g_threads = n < 1 ? 1 : n; return g_threads; }

// sizes: [ndof, numBlockColumn, numRemainedColumn, nBlockRowsEntries, nValues]
void oro_system_sizes(void *h, long long *out) {
    System *s = (System *)h;
    out[0] = s->ndof; out[1] = s->pat.numBlockColumn; out[2] = s->pat.numRemainedColumn;
    out[3] = (long long)s->pat.blockRows.size(); out[4] = s->pat.numEntries();
}
void oro_pattern(void *h, int *blockPcol, int *blockColumnOris, int *blockRows, int *elemMap) {
    System *s = (System *)h;
    std::memcpy(blockPcol, s->pat.blockPcol.data(), s->pat.blockPcol.size() * sizeof(int));
    std::memcpy(blockColumnOris, s->pat.blockColumnOris.data(), s->pat.blockColumnOris.size() * sizeof(int));
    std::memcpy(blockRows, s->pat.blockRows.data(), s->pat.blockRows.size() * sizeof(int));
    if (elemMap) std::memcpy(elemMap, s->elemMap.data(), s->elemMap.size() * sizeof(int));
}
int oro_global_index(void *h, int node, int axis) { return ((System *)h)->indexer.getGlobalIndex(node, axis); }

void oro_set_displacements(void *h, const double *u) {
    System *s = (System *)h;
    for (int i = 0; i < s->nNodes; i++) s->disp[i] = Vec3(u[3 * i], u[3 * i + 1], u[3 * i + 2]);
}
void oro_set_free_displacements(void *h, const double *d) { writeDisplacements(*(System *)h, d); }

void oro_mass(void *h, const double *rhoPerTet, double *M) { massMatrix(*(System *)h, rhoPerTet, M); }
void oro_body_force(void *h, const double *bodyPerTet, double *vws) { bodyForce(*(System *)h, bodyPerTet, vws); }
// forcer.cpp:35-50 (addSurfaceForceByNode), densified into vws
void oro_nodal_force(void *h, const double *f, int count, const int *nodes, double *vws) {
    System *s = (System *)h;
    for (int i = 0; i < count; i++)
        for (int j = 0; j < 3; j++)
            if (f[j] != 0.0) {
                int g = s->indexer.getGlobalIndex(nodes[i], j);
                if (g >= 0) vws[g] += f[j];
            }
}

// corotational plastic parameters per tet (kind 2); resets the plastic strain state to zero
void oro_set_plastic(void *h, const double *yield, const double *creep, const double *maxStrain) {
    System *s = (System *)h;
    s->plYield.assign(yield, yield + s->nTets);
    s->plCreep.assign(creep, creep + s->nTets);
    s->plMax.assign(maxStrain, maxStrain + s->nTets);
    s->plStrain.assign(9 * (size_t)s->nTets, 0.0);
}
void oro_get_plastic_strain(void *h, double *out) {
    System *s = (System *)h;
    std::copy(s->plStrain.begin(), s->plStrain.end(), out);
}

void oro_assemble(void *h, const int *kind, const double *E, const double *nu, double thrC, double thrN, double *K, double *vws) {
    System *s = (System *)h;
    Materials m = makeMaterials(s->nTets, kind, E, nu, thrC, thrN);
    assemble(*s, m, K, vws);
}

void oro_spmdv(void *h, const double *values, const double *x, double *y) { SpMDV(((System *)h)->pat, values, x, y); }
void oro_get_diagonal(void *h, const double *values, double *diag) { ((System *)h)->pat.getDiagonal(values, diag); }

int oro_cg(void *h, const double *values, const double *rhs, double *x, double tol, int maxIter, int pre, double p1, double p2,
           double *relres) {
    System *s = (System *)h;
    Preconditioner *P = makePre(pre, p1, p2);
    CG cg{tol, maxIter, P};
    int it = cg.solve(s->pat, values, rhs, x);
    if (relres) *relres = cg.lastRelRes;
    delete P;
    return it;
}

// Apply a preconditioner built from `values` once: result = M^-1 rhs
void oro_precond_apply(void *h, const double *values, int pre, double p1, double p2, const double *rhs, double *result) {
    System *s = (System *)h;
    Preconditioner *P = makePre(pre, p1, p2);
    P->Preprocess(s->pat, values);
    P->solve(s->ndof, rhs, result);
    delete P;
}

void oro_ic0_factor(void *h, const double *values, double pivotEps, double *Lcanonical) {
    System *s = (System *)h;
    IC0 ic(pivotEps);
    ic.Preprocess(s->pat, values);
    ic.exportCanonical(Lcanonical);
}

// Threshold IC factor as a dense lower-triangular matrix (small tests only)
void oro_ict_dense(void *h, const double *values, double eps, double relax, double *Ldense) {
    System *s = (System *)h;
    ThresholdIC ic(eps, relax);
    ic.Preprocess(s->pat, values);
    int n = s->ndof;
    std::memset(Ldense, 0, sizeof(double) * (size_t)n * n);
    for (int c = 0; c < n; c++)
        for (int q = ic.pcol[c]; q < ic.pcol[c + 1]; q++) Ldense[(size_t)ic.rows[q] * n + c] = ic.values[q];
}
// SAINV LDL^T factor: dense unit-lower L and D (small tests only)
void oro_ldlt_dense(void *h, const double *values, double se, double le, double *Ldense, double *D) {
    System *s = (System *)h;
    SainvLDLT f(se, le);
    f.Preprocess(s->pat, values);
    int n = s->ndof;
    std::memset(Ldense, 0, sizeof(double) * (size_t)n * n);
    for (int c = 0; c < n; c++) {
        Ldense[(size_t)c * n + c] = 1.0;
        if (c + 1 < (int)f.pcol.size())   // pcol has n entries: the last column has no sub-diagonal part
            for (int q = f.pcol[c]; q < f.pcol[c + 1]; q++) Ldense[(size_t)f.rows[q] * n + c] = f.values[q];
        D[c] = 1.0 / f.invDiagonal[c];
    }
}

// element-level entry points (unit tests)
void oro_polar(const double *F, double *R, double *S) { polarDecompostion3x3(F, R, S); }
// forceSemidefinite3x3 / forceSemidefinite2x2 (tetelement.cpp:446-499), in place
void oro_psd3(double *packed6) { forceSemidefinite3x3(packed6); }
void oro_psd2(double *diag, double *offdiag) { forceSemidefinite2x2(*diag, *offdiag); }
void oro_qr(const double *F, double *Q, double *R) { properQRDecompostion3x3(F, Q, R); }
void oro_eigsym(const double *tri, double *vals, double *vecs) { eigenSym3x3(tri, vals, vecs); }
void oro_corot_element(const double *X12, const double *x12, double E, double nu, double threshold, double *sub78, double *f12,
                       double *R9) {
    Vec3 ori[4], cur[4];
    for (int a = 0; a < 4; a++) { ori[a] = Vec3(X12[3 * a], X12[3 * a + 1], X12[3 * a + 2]); cur[a] = Vec3(x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]); }
    Scalar D[3], K0[90], DmInv[9], R[9];
    lameD(E, nu, D);
    corotPrecompute(ori, D, K0, DmInv);
    corotRotation(ori, cur, DmInv, threshold, R);
    corotSubStiffness(ori, cur, R, K0, sub78, f12);
    if (R9) std::memcpy(R9, R, sizeof(R));
}
void oro_corot_k0(const double *X12, double E, double nu, double *K0_90) {
    Vec3 ori[4];
    for (int a = 0; a < 4; a++) ori[a] = Vec3(X12[3 * a], X12[3 * a + 1], X12[3 * a + 2]);
    Scalar D[3], DmInv[9];
    lameD(E, nu, D);
    corotPrecompute(ori, D, K0_90, DmInv);
}
void oro_neo_element(const double *X12, const double *x12, double E, double nu, double threshold, double *sub78, double *f12) {
    Vec3 ori[4], cur[4];
    for (int a = 0; a < 4; a++) { ori[a] = Vec3(X12[3 * a], X12[3 * a + 1], X12[3 * a + 2]); cur[a] = Vec3(x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]); }
    Scalar lambda = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    for (int i = 0; i < 78; i++) sub78[i] = 0;
    neoElement(ori, cur, lambda, mu, threshold, sub78, f12);
}
void oro_neo_stress(const double *F, double E, double nu, double threshold, double *P, double *dPdF81) {
    Scalar lambda = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    Scalar diags[3], U[9], VT[9], invs[3], g[3], hh[6];
    modifiedDeformGradient(F, threshold, diags, U, VT);
    getInvariants(diags, invs);
    neoEnergyGradient(lambda, mu, invs, g);
    neoEnergyHessian(lambda, mu, invs, hh);
    piolaKirchhoff(diags, U, VT, invs, g, P);
    if (dPdF81) getdPdF(diags, g, hh, dPdF81);
}

// --- Newmark ---------------------------------------------------------------
void *oro_newmark_create(void *sys, const int *kind, const double *E, const double *nu, double thrC, double thrN,
                         const double *rhoPerTet, double beta, double gamma, double dt, double alphaM, double alphaK,
                         int pre, double p1, double p2, double tol, int maxIter, int initAccel) {
    System *s = (System *)sys;
    Newmark *nm = new Newmark();
    nm->s = s;
    nm->mat = makeMaterials(s->nTets, kind, E, nu, thrC, thrN);
    nm->pre = makePre(pre, p1, p2);
    nm->cg = CG{tol, maxIter, nm->pre};
    nm->beta = beta; nm->gamma = gamma; nm->h = dt; nm->alphaM = alphaM; nm->alphaK = alphaK;
    int n = s->ndof;
    size_t nv = (size_t)s->pat.numEntries();
    for (auto *vec : {&nm->d, &nm->v, &nm->a, &nm->fext, &nm->impulse, &nm->fint, &nm->rhs, &nm->tmp, &nm->dpred, &nm->vpred})
        vec->assign(n, 0.0);
    nm->M.assign(nv, 0.0); nm->K.assign(nv, 0.0); nm->T.assign(nv, 0.0);
    massMatrix(*s, rhoPerTet, nm->M.data());
    nm->initAccel = initAccel != 0;
    s->disp.assign(s->nNodes, Vec3());
    return nm;
}
void oro_newmark_destroy(void *h) { Newmark *nm = (Newmark *)h; delete nm->pre; delete nm; }
void oro_newmark_set_external(void *h, const double *f) { Newmark *nm = (Newmark *)h; std::memcpy(nm->fext.data(), f, nm->fext.size() * sizeof(double)); }
void oro_newmark_set_impulse(void *h, const double *f) { Newmark *nm = (Newmark *)h; std::memcpy(nm->impulse.data(), f, nm->impulse.size() * sizeof(double)); }
void oro_newmark_set_state(void *h, const double *d, const double *v, const double *a, int steps) {
    Newmark *nm = (Newmark *)h;
    size_t n = nm->d.size();
    std::memcpy(nm->d.data(), d, n * 8); std::memcpy(nm->v.data(), v, n * 8); std::memcpy(nm->a.data(), a, n * 8);
    nm->steps = steps;
    writeDisplacements(*nm->s, d);
}
// runs nSteps; iters[k] = PCG iterations of step k (nullable); returns init-accel iterations
int oro_newmark_step(void *h, int nSteps, int *iters, double *relres) {
    Newmark *nm = (Newmark *)h;
    for (int k = 0; k < nSteps; k++) {
        nm->step();
        if (iters) iters[k] = nm->lastIters;
        if (relres) relres[k] = nm->lastRelRes;
    }
    return nm->initIters;
}
void oro_newmark_get_state(void *h, double *d, double *v, double *a) {
    Newmark *nm = (Newmark *)h;
    size_t n = nm->d.size();
    if (d) std::memcpy(d, nm->d.data(), n * 8);
    if (v) std::memcpy(v, nm->v.data(), n * 8);
    if (a) std::memcpy(a, nm->a.data(), n * 8);
}
void oro_newmark_mass(void *h, double *M) { Newmark *nm = (Newmark *)h; std::memcpy(M, nm->M.data(), nm->M.size() * 8); }

// --- Backward Euler (the reference's own full-order integrator) --------------
void *oro_be_create(void *sys, const int *kind, const double *E, const double *nu, double thrC, double thrN,
                    const double *rhoPerTet, double dt, double alphaM, double alphaK, int pre, double p1, double p2,
                    double tol, int maxIter) {
    System *s = (System *)sys;
    BackwardEuler *be = new BackwardEuler();
    be->s = s;
    be->mat = makeMaterials(s->nTets, kind, E, nu, thrC, thrN);
    be->pre = makePre(pre, p1, p2);
    be->cg = CG{tol, maxIter, be->pre};
    be->h = dt; be->alphaM = alphaM; be->alphaK = alphaK;
    int n = s->ndof;
    size_t nv = (size_t)s->pat.numEntries();
    for (auto *vec : {&be->v, &be->dv, &be->fint, &be->fext, &be->rhs}) vec->assign(n, 0.0);
    be->K.assign(nv, 0.0); be->M.assign(nv, 0.0); be->D.assign(nv, 0.0);
    massMatrix(*s, rhoPerTet, be->M.data());
    s->disp.assign(s->nNodes, Vec3());
    return be;
}
void oro_be_destroy(void *h) { BackwardEuler *be = (BackwardEuler *)h; delete be->pre; delete be; }
void oro_be_set_external(void *h, const double *f) { BackwardEuler *be = (BackwardEuler *)h; std::memcpy(be->fext.data(), f, be->fext.size() * 8); }
int oro_be_step(void *h) { BackwardEuler *be = (BackwardEuler *)h; be->step(); return be->cg.lastIters; }
void oro_be_update_mesh(void *h) { ((BackwardEuler *)h)->updateMesh(); }
void oro_be_get_velocity(void *h, double *v) { BackwardEuler *be = (BackwardEuler *)h; std::memcpy(v, be->v.data(), be->v.size() * 8); }
// MeshPlaneCollisionDampingForcer (penaltyDampingForcer.cpp:8-61) for nPairs (plane, tet)
// pairs in order; planes: 4 per pair (normal, constant); velNodes: 3 per mesh node
void oro_plane_damping(void *sys, int nPairs, const double *planes, const int *tets, double coeff,
                       const double *velNodes, double *values) {
    System *s = (System *)sys;
    std::vector<Vec3> vel(s->nNodes);
    for (int i = 0; i < s->nNodes; i++) vel[i] = Vec3(velNodes[3 * i], velNodes[3 * i + 1], velNodes[3 * i + 2]);
    for (int k = 0; k < nPairs; k++) {
        Plane pl{Vec3(planes[4 * k], planes[4 * k + 1], planes[4 * k + 2]), planes[4 * k + 3]};
        planeDampingPair(*s, pl, tets[k], coeff, vel.data(), values);
    }
}
// addExternalDampingForceMatrix with the forcer reading the integrator's own velocities
// (getMeshVerticesVelocities) and the mesh displacements
void oro_be_add_plane_damping(void *h, int nPairs, const double *planes, const int *tets, double coeff) {
    BackwardEuler *be = (BackwardEuler *)h;
    std::vector<Vec3> vel = nodeVelocities(*be->s, be->v.data());
    for (int k = 0; k < nPairs; k++) {
        Plane pl{Vec3(planes[4 * k], planes[4 * k + 1], planes[4 * k + 2]), planes[4 * k + 3]};
        planeDampingPair(*be->s, pl, tets[k], coeff, vel.data(), be->D.data());
    }
}
void oro_be_clear_damping(void *h) { BackwardEuler *be = (BackwardEuler *)h; std::fill(be->D.begin(), be->D.end(), 0.0); }
void oro_be_damping(void *h, double *D) { BackwardEuler *be = (BackwardEuler *)h; std::memcpy(D, be->D.data(), be->D.size() * 8); }
void oro_be_set_velocity(void *h, const double *v) { BackwardEuler *be = (BackwardEuler *)h; std::memcpy(be->v.data(), v, be->v.size() * 8); }
void oro_plane_tet_intersection(const double *plane4, const double *pts12, double *centroid, double *dir) {
    Plane pl{Vec3(plane4[0], plane4[1], plane4[2]), plane4[3]};
    Vec3 p[4];
    for (int i = 0; i < 4; i++) p[i] = Vec3(pts12[3 * i], pts12[3 * i + 1], pts12[3 * i + 2]);
    Vec3 c, d;
    planeTetIntersection(pl, p, c, d);
    for (int i = 0; i < 3; i++) { centroid[i] = c[i]; dir[i] = d[i]; }
}
void oro_barycentric(const double *pts12, const double *q, double *coords) {
    Vec3 p[4];
    for (int i = 0; i < 4; i++) p[i] = Vec3(pts12[3 * i], pts12[3 * i + 1], pts12[3 * i + 2]);
    barycentric(p, Vec3(q[0], q[1], q[2]), coords);
}
// MeshRelabeler::getNewLables(elementCount, nodePerElementCount, newNodeLables, elements)
// (meshRelabler.cpp:468-498): new label per node, and the elements relabelled and sorted by
// the sum of their labels.  Returns -1 for a disconnected mesh.
int oro_gps(int nNodes, int nElems, int npe, int *elements, int *labels) {
    GpsGraph g;
    g.adj.resize(nNodes);
    std::set<std::pair<int, int>> seen;
    for (int e = 0; e < nElems; e++) {
        const int *el = elements + (size_t)e * npe;
        for (int j = 0; j < npe; j++)
            for (int k = j + 1; k < npe; k++) {
                int a = std::min(el[j], el[k]), b = std::max(el[j], el[k]);
                if (a == b || !seen.insert({a, b}).second) continue;
                g.adj[a].push_front(b);
                g.adj[b].push_front(a);
            }
    }
    std::vector<int> lab;
    if (!gpsLabels(g, lab)) return -1;
    std::copy(lab.begin(), lab.end(), labels);
    std::vector<std::pair<long long, int>> order(nElems);
    std::vector<int> relab((size_t)nElems * npe);
    for (int e = 0; e < nElems; e++) {
        long long score = 0;
        for (int j = 0; j < npe; j++) {
            relab[(size_t)e * npe + j] = lab[elements[(size_t)e * npe + j]];
            score += relab[(size_t)e * npe + j];
        }
        order[e] = {score, e};
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<long long, int> &a, const std::pair<long long, int> &b) { return a.first < b.first; });
    for (int i = 0; i < nElems; i++)
        std::memcpy(elements + (size_t)i * npe, &relab[(size_t)order[i].second * npe], sizeof(int) * npe);
    return 0;
}

void oro_get_node_displacements(void *sys, double *u) {
    System *s = (System *)sys;
    for (int i = 0; i < s->nNodes; i++) { u[3 * i] = s->disp[i].x; u[3 * i + 1] = s->disp[i].y; u[3 * i + 2] = s->disp[i].z; }
}

}  // extern "C"
