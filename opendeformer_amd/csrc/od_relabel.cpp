// Gibbs-Poole-Stockmeyer node relabelling of a mesh graph (host code): the algorithm of
// MeshRelabeler::getNewLables (core/meshRelabler.cpp:57-530, core/meshRelabeler.h:16-135).
//
// Orders the reference leaves to its standard library are fixed as follows: edges are met
// in first-insertion order (element order, node pairs j < k); arc lists grow at the front;
// level sorts by degree, the sort of components by size and the sort of elements by label
// sum are stable; the max-degree queue of processSigleVert pops equal degrees in push order.
//
// The structures are sized for large meshes: arc lists are vectors kept in reverse (the
// list front is the vector back, so a prepend is a push_back), and the level lists of the
// numbering phase are intrusive doubly linked lists with a level tag per vertex, so the
// membership tests and removals the reference does with std::find are O(1).
#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <queue>
#include <string>
#include <unordered_set>
#include <vector>

#include "od_setup.h"

namespace od {
namespace {

struct Graph {
    std::vector<std::vector<int>> rev;   // arc list of v, last element = list front
    int degree(int v) const { return (int)rev[v].size(); }
    template <typename F>
    void forArcs(int v, F f) const {     // front to back
        const std::vector<int> &a = rev[v];
        for (size_t i = a.size(); i-- > 0;) f(a[i]);
    }
    void prepend(int v, int u) { rev[v].push_back(u); }
    void eraseFirst(int v, int u) {      // first occurrence from the front (MeshGraphVertexNode::deleteArc)
        std::vector<int> &a = rev[v];
        for (size_t i = a.size(); i-- > 0;)
            if (a[i] == u) { a.erase(a.begin() + (std::ptrdiff_t)i); return; }
    }
};

// RootedLevelStructure::generateLevels (meshRelabler.cpp:57-95)
struct Levels {
    int root = -1;
    std::vector<int> verts;               // level 1, level 2, ... concatenated
    std::vector<int> start;               // start[i] .. start[i+1]: level i+1
    int levelSize() const { return (int)start.size(); }   // levels + the root's
    int width() const {
        int w = INT_MIN;
        for (size_t i = 0; i + 1 < start.size(); i++) w = std::max(w, start[i + 1] - start[i]);
        return w;
    }
    void generate(const Graph &g, int r, std::vector<int> &mark, int stamp) {
        root = r;
        verts.clear();
        start.assign(1, 0);
        mark[r] = stamp;
        g.forArcs(r, [&](int u) { verts.push_back(u); mark[u] = stamp; });
        size_t lo = 0;
        while (lo < verts.size()) {
            start.push_back((int)verts.size());
            const size_t hi = verts.size();
            for (size_t i = lo; i < hi; i++)
                g.forArcs(verts[i], [&](int u) {
                    if (mark[u] != stamp) { mark[u] = stamp; verts.push_back(u); }
                });
            lo = hi;
        }
        // start now has one entry per non-empty level plus the trailing end
    }
};

// Level lists of the numbering phase.
struct LevelLists {
    std::vector<int> prev, next, tag, head, tail;
    LevelLists(int n, int nLevels) : prev(n, -1), next(n, -1), tag(n, -1), head(nLevels, -1), tail(nLevels, -1) {}
    bool in(int v, int l) const { return tag[v] == l; }
    bool empty(int l) const { return head[l] < 0; }
    void pushBack(int l, int v) {
        tag[v] = l; prev[v] = tail[l]; next[v] = -1;
        if (tail[l] >= 0) next[tail[l]] = v; else head[l] = v;
        tail[l] = v;
    }
    void unlink(int v) {
        const int l = tag[v];
        if (l < 0) return;
        if (prev[v] >= 0) next[prev[v]] = next[v]; else head[l] = next[v];
        if (next[v] >= 0) prev[next[v]] = prev[v]; else tail[l] = prev[v];
        tag[v] = -1; prev[v] = next[v] = -1;
    }
    int popFront(int l) { const int v = head[l]; unlink(v); return v; }
};

// MeshRelabeler::processSigleVert (meshRelabler.cpp:148-172)
void processVert(const Graph &g, int v, int level, LevelLists &L, std::queue<int> &working, std::vector<int> &label,
                 int &next, std::vector<int> &toBe) {
    label[v] = next++;
    working.push(v);
    toBe.clear();
    g.forArcs(v, [&](int u) {
        if (L.in(u, level)) { toBe.push_back(u); L.unlink(u); }
    });
    // max degree first, equal degrees in push order
    std::stable_sort(toBe.begin(), toBe.end(), [&](int a, int b) { return g.degree(a) > g.degree(b); });
    for (int u : toBe) { label[u] = next++; working.push(u); }
}

// MeshRelabeler::setNewNodeLables (meshRelabler.cpp:174-466)
bool gpsLabels(Graph &g, std::vector<int> &out) {
    const int n = (int)g.rev.size();
    std::vector<int> mark(n, 0);
    int stamp = 1;
    int s0 = 0;
    for (int v = 1; v < n; v++)
        if (g.degree(v) < g.degree(s0)) s0 = v;
    Levels startL, endL, mayL;
    startL.generate(g, s0, mark, stamp++);
    if ((int)startL.verts.size() + 1 != n) return false;
    bool found = false;
    while (!found) {
        found = true;
        int endWidth = INT_MAX;
        const int last = startL.levelSize() - 1;
        std::vector<int> cand(startL.verts.begin() + startL.start[last - 1], startL.verts.begin() + startL.start[last]);
        std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return g.degree(a) < g.degree(b); });
        std::copy(cand.begin(), cand.end(), startL.verts.begin() + startL.start[last - 1]);
        for (int c : cand) {
            mayL.generate(g, c, mark, stamp++);
            if (mayL.levelSize() > startL.levelSize()) {
                found = false;
                std::swap(startL, mayL);
                break;
            }
            const int w = mayL.width();
            if (w < endWidth) { std::swap(endL, mayL); endWidth = w; }
        }
    }
    // level pairs and the vertices both structures put on the same level (:224-283)
    const int startSize = startL.levelSize(), endSize = endL.levelSize();
    const int size = std::max(startSize, endSize);
    const bool startSmaller = endL.width() > startL.width();
    std::vector<int> sLev(n, 0), eLev(n, INT_MIN);
    for (int i = 1; i < startSize; i++)
        for (int k = startL.start[i - 1]; k < startL.start[i]; k++) sLev[startL.verts[k]] = i;
    std::vector<int> deleted, remained;
    std::vector<std::vector<int>> newLev(size);
    auto place = [&](int v, int e) {
        eLev[v] = e;
        if (sLev[v] == e) { deleted.push_back(v); newLev[e].push_back(v); }
        else remained.push_back(v);
    };
    place(endL.root, endSize - 1);
    for (int i = 1; i < endSize; i++)
        for (int k = endL.start[i - 1]; k < endL.start[i]; k++) place(endL.verts[k], endSize - 1 - i);
    // deleteArc (:285-293)
    for (int v : deleted) {
        std::vector<int> arcs;
        g.forArcs(v, [&](int u) { arcs.push_back(u); });
        for (int u : arcs) g.eraseFirst(u, v);
    }
    // connected components of the remaining vertices, in the reference's queue order (:295-332)
    std::vector<std::vector<int>> comps;
    {
        const int st = stamp++;
        std::queue<int> working;
        for (int v : remained) {
            if (mark[v] == st) continue;
            comps.emplace_back();
            mark[v] = st;
            comps.back().push_back(v);
            g.forArcs(v, [&](int u) { working.push(u); });
            while (!working.empty()) {
                const int a = working.front();
                working.pop();
                if (mark[a] == st) continue;
                mark[a] = st;
                comps.back().push_back(a);
                g.forArcs(a, [&](int u) { working.push(u); });
            }
        }
    }
    std::stable_sort(comps.begin(), comps.end(),
                     [](const std::vector<int> &a, const std::vector<int> &b) { return a.size() < b.size(); });
    bool selectSecond = false;
    std::vector<int> c1(size), c2(size);
    for (size_t i = 0; i < comps.size(); i++) {
        std::fill(c1.begin(), c1.end(), 0);
        std::fill(c2.begin(), c2.end(), 0);
        for (int v : comps[i]) { c1[sLev[v]]++; c2[eLev[v]]++; }
        const int m1 = *std::max_element(c1.begin(), c1.end()), m2 = *std::max_element(c2.begin(), c2.end());
        if (i == 0 && (m1 > m2 || (m1 == m2 && !startSmaller))) selectSecond = true;
        const bool useStart = m1 < m2 ? true : (m1 > m2 ? false : startSmaller);
        for (int v : comps[i]) newLev[useStart ? sLev[v] : eLev[v]].push_back(v);
    }
    // insertArc restores the arcs at the list fronts (:375-383)
    for (int v : deleted) {
        std::vector<int> arcs;
        g.forArcs(v, [&](int u) { arcs.push_back(u); });
        for (int u : arcs) g.prepend(u, v);
    }
    LevelLists L(n, size);
    for (int l = 0; l < size; l++) {
        std::stable_sort(newLev[l].begin(), newLev[l].end(), [&](int a, int b) { return g.degree(a) < g.degree(b); });
        for (int v : newLev[l]) L.pushBack(l, v);
        std::vector<int>().swap(newLev[l]);
    }
    // numbering (:389-457)
    bool interchange = false;
    int root = startL.root, init = 0, pace = 1;
    if (g.degree(startL.root) > g.degree(endL.root)) { interchange = true; root = endL.root; init = size - 1; pace = -1; }
    L.unlink(root);
    std::vector<int> label(n, -1), toBe;
    int next = 0;
    std::queue<int> working;
    processVert(g, root, init, L, working, label, next, toBe);
    while (!L.empty(init)) processVert(g, L.popFront(init), init, L, working, label, next, toBe);
    for (int level = init + pace; level >= 0 && level < size; level += pace) {
        int process = -1;
        bool unfound = true;
        while (unfound && !working.empty()) {
            const int v = working.front();
            g.forArcs(v, [&](int u) {
                if (L.in(u, level)) {
                    unfound = false;
                    if (process < 0 || g.degree(process) > g.degree(u)) process = u;
                }
            });
            working.pop();
        }
        std::queue<int>().swap(working);
        if (process < 0) {                   // a level no earlier vertex reaches (not in a connected mesh)
            if (L.empty(level)) continue;
            process = L.head[level];
        }
        L.unlink(process);
        processVert(g, process, level, L, working, label, next, toBe);
        while (!L.empty(level)) processVert(g, L.popFront(level), level, L, working, label, next, toBe);
    }
    out.resize(n);
    for (int v = 0; v < n; v++) out[v] = (interchange != selectSecond) ? label[v] : (n - 1) - label[v];
    return true;
}

}  // namespace

const char *relabelGps(int nNodes, int nElems, int npe, int *elements, int *labels) {
    if (nNodes <= 0 || nElems <= 0 || npe < 2 || !elements || !labels) return "bad arguments";
    for (int64_t k = 0; k < (int64_t)nElems * npe; k++)
        if (elements[k] < 0 || elements[k] >= nNodes) return "element names a node outside the mesh";
    Graph g;
    g.rev.resize(nNodes);
    {
        // edges in first-insertion order (generateGraphFromElementIndices, meshRelabler.cpp:127-146)
        std::unordered_set<uint64_t> seen;
        seen.reserve((size_t)nElems * 4);
        for (int e = 0; e < nElems; e++) {
            const int *el = elements + (size_t)e * npe;
            for (int j = 0; j < npe; j++)
                for (int k = j + 1; k < npe; k++) {
                    const int a = std::min(el[j], el[k]), b = std::max(el[j], el[k]);
                    if (a == b || !seen.insert(((uint64_t)a << 32) | (uint32_t)b).second) continue;
                    g.prepend(a, b);
                    g.prepend(b, a);
                }
        }
    }
    std::vector<int> lab;
    if (!gpsLabels(g, lab)) return "the mesh graph is not connected";
    std::memcpy(labels, lab.data(), sizeof(int) * (size_t)nNodes);
    // elements relabelled and sorted by label sum, stable (meshRelabler.cpp:474-497)
    std::vector<int> relab((size_t)nElems * npe);
    std::vector<int64_t> score(nElems, 0);
    for (int e = 0; e < nElems; e++)
        for (int j = 0; j < npe; j++) {
            relab[(size_t)e * npe + j] = lab[elements[(size_t)e * npe + j]];
            score[e] += relab[(size_t)e * npe + j];
        }
    std::vector<int> order(nElems);
    for (int e = 0; e < nElems; e++) order[e] = e;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return score[a] < score[b]; });
    for (int i = 0; i < nElems; i++)
        std::memcpy(elements + (size_t)i * npe, &relab[(size_t)order[i] * npe], sizeof(int) * npe);
    return nullptr;
}

}  // namespace od
