// Host factorisations of the reference's scalar preconditioners and the level schedules of
// their device sweeps (od_precond.h).  Every product and sum keeps the reference's operand
// order, so the factors round as the reference's do.
#if defined(__clang__)
#pragma clang fp contract(off)
#endif

#include <algorithm>
#include <cmath>
#include <utility>

#include "od_precond.h"

namespace od {
namespace {

// Packed diagonal block: column offset o starts at entry kDiagOff[o] (core/sparseMatrix.h:85-88).
constexpr int kDiagOff[3] = {0, 3, 5};

// Accumulator indexed by row with first-insertion order (the reference's FastSparseVector,
// core/latool.h:453-512): add() sums into an existing entry or appends a new one.
struct OrderedAccumulator {
    std::vector<int> slot;       // 0: absent, else 1 + position
    std::vector<int> idx;
    std::vector<double> val;
    explicit OrderedAccumulator(int n) : slot((size_t)n, 0) {}
    void clear() {
        for (int i : idx) slot[i] = 0;
        idx.clear();
        val.clear();
    }
    void append(int i, double v) {
        slot[i] = (int)idx.size() + 1;
        idx.push_back(i);
        val.push_back(v);
    }
    void add(int i, double v) {
        const int p = slot[i];
        if (p) val[p - 1] += v;
        else append(i, v);
    }
    double at(int i) const {
        const int p = slot[i];
        return p ? val[p - 1] : 0.0;
    }
};

// Column c = 3R + o of the scalar lower triangle (getColumn, core/sparseMatrix.h:297-365): rows
// c .. 3R+2 of the diagonal block, then rows 3J .. 3J+2 of each off-diagonal block (R, J) in
// increasing J (entry 3 o + i of the block), exact zeros skipped.
void lowerColumn(const Topology &t, const double *vals, int c, OrderedAccumulator &w) {
    w.clear();
    const int R = c / 3, o = c % 3;
    const double *d = vals + t.valOff[R] + kDiagOff[o];
    for (int i = 0; i < 3 - o; i++)
        if (d[i] != 0.0) w.append(c + i, d[i]);
    for (int b = t.rowPtr[R] + 1; b < t.rowPtr[R + 1]; b++) {
        const double *v = vals + t.blkOff[b] + 3 * o;
        const int row = 3 * t.colIdx[b];
        for (int i = 0; i < 3; i++)
            if (v[i] != 0.0) w.append(row + i, v[i]);
    }
}

double diagonalEntry(const Topology &t, const double *vals, int c) { return vals[t.valOff[c / 3] + kDiagOff[c % 3]]; }

}  // namespace

void thresholdIcFactor(const Topology &t, const double *vals, double eps, double relaxedScale, ScalarFactor &L) {
    const int n = t.ndof;
    L.n = n;
    L.pcol.assign(1, 0);
    L.rows.clear();
    L.values.clear();
    L.invDiag.clear();
    if (n == 0) return;
    std::vector<double> diags((size_t)n);
    for (int c = 0; c < n; c++) diags[c] = diagonalEntry(t, vals, c);
    // head[r]: (column, position of its entry in row r) of the first column whose next
    // unconsumed entry lies in row r; the columns waiting on row r chain through next[col]
    std::vector<std::pair<int, int>> head((size_t)n, {-1, -1}), next((size_t)n, {-1, -1});
    OrderedAccumulator w(n);
    std::vector<std::pair<int, double>> kept;
    const double eps2 = eps * eps;

    auto finishColumn = [&](int c) {
        // the first entry is the diagonal; the running diagonal lives in diags
        double diag = diags[c];
        kept.clear();
        for (size_t q = 1; q < w.idx.size(); q++) {
            const int row = w.idx[q];
            const double v = w.val[q];
            if (v == 0.0) continue;
            if (v * v < std::fabs(eps2 * diag * diags[row])) {
                const double comp = std::fabs(v) * relaxedScale;   // dropped: compensate both diagonals
                diag += comp;
                diags[row] += comp;
            } else {
                kept.emplace_back(row, v);
            }
        }
        const int start = L.pcol[c];
        L.pcol.push_back(start + 1 + (int)kept.size());
        std::sort(kept.begin(), kept.end(),
                  [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
        if (!kept.empty()) {
            const int r = kept[0].first;
            next[c] = head[r];
            head[r] = {c, start + 1};
        }
        if (diag <= 0.0) diag = eps2 * diag * diag;
        const double root = std::sqrt(diag);
        L.values.push_back(root);
        L.rows.push_back(c);
        const double inv = 1.0 / root;
        for (const auto &e : kept) {
            const double v = e.second * inv;
            L.rows.push_back(e.first);
            L.values.push_back(e.second * inv);
            diags[e.first] -= v * v;
        }
    };

    lowerColumn(t, vals, 0, w);
    finishColumn(0);
    for (int i = 1; i < n - 1; i++) {
        lowerColumn(t, vals, i, w);
        std::pair<int, int> cur = head[i];
        while (cur.first >= 0) {
            const int j = cur.first, at = cur.second, after = at + 1;
            const double upper = L.values[at];
            for (int k = after; k < L.pcol[j + 1]; k++) w.add(L.rows[k], -upper * L.values[k]);
            cur = next[j];
            if (after < L.pcol[j + 1]) {
                // column j waits next on the row of its following entry
                const int r = L.rows[after];
                next[j] = head[r];
                head[r] = {j, after};
            }
        }
        finishColumn(i);
    }
    if (n > 1) {
        // the last column: no pivot fix (inCholeskyPreconditioner.cpp:124-126)
        L.values.push_back(std::sqrt(diags[n - 1]));
        L.rows.push_back(n - 1);
        L.pcol.push_back(L.pcol[n - 1] + 1);
    }
}

void sainvLdltFactor(const Topology &t, const double *vals, double sainvEps, double ldltEps, ScalarFactor &L) {
    const int n = t.ndof;
    L.n = n;
    L.pcol.assign(1, 0);
    L.rows.clear();
    L.values.clear();
    L.invDiag.assign((size_t)n, 0.0);
    if (n == 0) return;
    // Full symmetric column lists (getFullIndices, core/sparseMatrix.h:454-526): for each
    // column, (row, value index) over both triangles, in the order the reference builds them.
    std::vector<std::vector<std::pair<int, int>>> full((size_t)n);
    for (int R = 0; R < t.nRows; R++) {
        const int c = 3 * R;
        int idx = t.valOff[R];
        for (int i = 0; i < 3; i++) {
            full[c + i].emplace_back(c + i, idx++);
            for (int j = i + 1; j < 3; j++) {
                full[c + i].emplace_back(c + j, idx);
                full[c + j].emplace_back(c + i, idx);
                idx++;
            }
        }
        for (int b = t.rowPtr[R] + 1; b < t.rowPtr[R + 1]; b++) {
            const int row = 3 * t.colIdx[b];
            idx = t.blkOff[b];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    full[c + i].emplace_back(row + j, idx);
                    full[row + j].emplace_back(c + i, idx);
                    idx++;
                }
        }
    }
    // Inverse-factor entries: node k holds (row, col, value); the entries of each row form a
    // doubly linked ring through sentinel node `row` (new entries at its front), and col's
    // entries are listed in colNodes[col] in insertion order.
    struct Node { int prev, next, row, col; double value; };
    std::vector<Node> nodes((size_t)n);
    std::vector<int> freeNodes;
    auto newNode = [&](int row, int col, double v) {
        int k;
        if (!freeNodes.empty()) { k = freeNodes.back(); freeNodes.pop_back(); }
        else { k = (int)nodes.size(); nodes.push_back(Node{}); }
        nodes[k].row = row; nodes[k].col = col; nodes[k].value = v;
        Node &h = nodes[row];
        nodes[k].next = h.next;
        nodes[k].prev = row;
        nodes[h.next].prev = k;
        h.next = k;
        return k;
    };
    auto unlink = [&](int k) {
        nodes[nodes[k].prev].next = nodes[k].next;
        nodes[nodes[k].next].prev = nodes[k].prev;
        freeNodes.push_back(k);
    };
    std::vector<std::vector<int>> colNodes((size_t)n);
    for (int i = 0; i < n; i++) { nodes[i].prev = nodes[i].next = i; nodes[i].row = i; nodes[i].col = -1; }
    for (int i = 0; i < n; i++) colNodes[i].push_back(newNode(i, i, 1.0));

    OrderedAccumulator vec(n), lower(n);
    std::vector<int> zIdx;
    std::vector<double> zVal;
    auto takeColumn = [&](int c) {   // moves column c of the inverse factor into (zIdx, zVal)
        zIdx.clear();
        zVal.clear();
        for (int k : colNodes[c]) {
            zIdx.push_back(nodes[k].row);
            zVal.push_back(nodes[k].value);
            unlink(k);
        }
        colNodes[c].clear();
        colNodes[c].shrink_to_fit();
    };
    auto productWithA = [&]() {      // vec = A z (SpMSV, core/numerMethod.h:282-300)
        for (size_t q = 0; q < zIdx.size(); q++) {
            const double e = zVal[q];
            for (const auto &pr : full[zIdx[q]]) vec.add(pr.first, vals[pr.second] * e);
        }
    };
    auto dotZ = [&]() {              // z . vec in z's order (SparseVector::operator*, core/latool.h:594-600)
        double dot = 0.0;
        for (size_t q = 0; q < zIdx.size(); q++) dot += zVal[q] * vec.at(zIdx[q]);
        return dot;
    };

    int count = 0;
    for (int i = 0; i < n - 1; i++) {
        takeColumn(i);
        productWithA();
        const double inv = 1.0 / dotZ();
        L.invDiag[i] = inv;
        // lower[j] = (A z_i) . z_j over the later columns j that share a row with A z_i
        for (size_t q = 0; q < vec.idx.size(); q++) {
            const double value = vec.val[q];
            if (value == 0.0) continue;
            const int row = vec.idx[q];
            for (int k = nodes[row].next; k != row; k = nodes[k].next) lower.add(nodes[k].col, value * nodes[k].value);
        }
        for (size_t q = 0; q < lower.idx.size(); q++) {
            const double l = lower.val[q] * inv;
            if (std::fabs(l) > ldltEps) {
                L.values.push_back(l);
                L.rows.push_back(lower.idx[q]);
                ++count;
            }
        }
        L.pcol.push_back(count);
        vec.clear();
        // z_j -= (lower_j / d_i) z_i, entries above sainvEps kept
        for (size_t q = 0; q < lower.idx.size(); q++) {
            const int j = lower.idx[q];
            double factor = lower.val[q];
            if (factor == 0.0) continue;
            factor *= inv;
            for (int k : colNodes[j]) {
                vec.append(nodes[k].row, nodes[k].value);
                unlink(k);
            }
            colNodes[j].clear();
            for (size_t p = 0; p < zIdx.size(); p++) vec.add(zIdx[p], -factor * zVal[p]);
            for (size_t p = 0; p < vec.idx.size(); p++) {
                const double v = vec.val[p];
                if (std::fabs(v) > sainvEps) colNodes[j].push_back(newNode(vec.idx[p], j, v));
            }
            vec.clear();
        }
        lower.clear();
    }
    takeColumn(n - 1);
    productWithA();
    L.invDiag[n - 1] = 1.0 / dotZ();
    vec.clear();
    L.pcol.push_back(count);   // the last column holds no entries
}

namespace {
// Levels: forward level of row i = 1 + the largest level among the rows it reads (0 if none);
// runs of levels narrower than kScalarWideLevel share one single-workgroup launch.
void schedule(int n, const std::vector<int> &ptr, const std::vector<int> &dep, bool descending, ScalarSweepPlan &p) {
    std::vector<int> lev((size_t)n, 0);
    int nl = 0;
    for (int s = 0; s < n; s++) {
        const int i = descending ? n - 1 - s : s;
        int l = 0;
        for (int e = ptr[i]; e < ptr[i + 1]; e++) l = std::max(l, lev[dep[e]] + 1);
        lev[i] = l;
        nl = std::max(nl, l + 1);
    }
    p.levelPtr.assign((size_t)nl + 1, 0);
    for (int i = 0; i < n; i++) p.levelPtr[lev[i] + 1]++;
    for (int l = 0; l < nl; l++) p.levelPtr[l + 1] += p.levelPtr[l];
    p.levelRows.assign((size_t)n, 0);
    std::vector<int> fill(p.levelPtr.begin(), p.levelPtr.end() - 1);
    for (int i = 0; i < n; i++) p.levelRows[fill[lev[i]]++] = i;
    p.launches.clear();
    for (int l = 0; l < nl;) {
        if (p.levelPtr[l + 1] - p.levelPtr[l] >= kScalarWideLevel) {
            p.launches.push_back({l, l + 1, true});
            l++;
            continue;
        }
        int e = l;
        while (e < nl && p.levelPtr[e + 1] - p.levelPtr[e] < kScalarWideLevel) e++;
        p.launches.push_back({l, e, false});
        l = e;
    }
}
}  // namespace

void buildScalarSweeps(const ScalarFactor &L, bool unit, ScalarSweepsHost &s) {
    const int n = L.n;
    s.n = n;
    s.unit = unit;
    const int firstOff = unit ? 0 : 1;   // IC columns store their diagonal first
    // backward: column i's strictly lower entries in stored order
    s.bPtr.assign((size_t)n + 1, 0);
    s.bRow.clear();
    s.bVal.clear();
    s.diag.assign((size_t)n, 0.0);
    for (int c = 0; c < n; c++) {
        for (int e = L.pcol[c] + firstOff; e < L.pcol[c + 1]; e++) {
            s.bRow.push_back(L.rows[e]);
            s.bVal.push_back(L.values[e]);
        }
        s.bPtr[c + 1] = (int)s.bRow.size();
        s.diag[c] = unit ? L.invDiag[c] : L.values[L.pcol[c]];
    }
    // forward: row i's entries L(i, k), k < i, in increasing k
    s.fPtr.assign((size_t)n + 1, 0);
    for (int e = 0; e < (int)s.bRow.size(); e++) s.fPtr[s.bRow[e] + 1]++;
    for (int i = 0; i < n; i++) s.fPtr[i + 1] += s.fPtr[i];
    s.fCol.assign(s.bRow.size(), 0);
    s.fVal.assign(s.bRow.size(), 0.0);
    std::vector<int> fill(s.fPtr.begin(), s.fPtr.end() - 1);
    for (int c = 0; c < n; c++)
        for (int e = s.bPtr[c]; e < s.bPtr[c + 1]; e++) {
            const int q = fill[s.bRow[e]]++;
            s.fCol[q] = c;
            s.fVal[q] = s.bVal[e];
        }
    schedule(n, s.fPtr, s.fCol, false, s.fwd);
    schedule(n, s.bPtr, s.bRow, true, s.bwd);
}

}  // namespace od
