// Host-side set-up of the device-resident mesh: DOF indexing, the canonical
// blocked-symmetric pattern, the per-element 78-entry map, the block-level
// assembly map, the transposed block index used by the SpMV and the IC(0)
// sweeps, the IC(0) level schedule and the per-instance row chunks.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace od {

struct Topology {
    // ---- inputs -------------------------------------------------------------
    int nNodes = 0, nTets = 0;
    std::vector<double> X;              // 3 per node
    std::vector<int> tets;              // 4 per tet
    std::vector<int> constrained;       // sorted constrained DOF ids (3*node+axis)
    std::vector<int> instNodeStart;     // instance node ranges (size nInst+1)
    std::vector<int> instTetStart;      // instance tet ranges (size nInst+1)

    // ---- DOF indexing (core/nodeIndexer.cpp:12-25 semantics) ----------------
    // Block row R holds the free DOFs 3R .. 3R+2 of the reference's numbering.  With whole-node
    // constraints these are the three DOFs of one free node.  When some node is partially
    // constrained (`partial`) a block row is a DOF triple that may span nodes, and the last
    // row is padded to three DOFs when ndofFree is not a multiple of 3: the kernels' vectors
    // then hold ndof = 3 * nRows entries, the padding entries stay zero, and the matrix holds
    // 1.0 on their diagonal and zeros elsewhere, so the padded system decouples them.
    int ndof = 0;                       // DOF entries of the kernels' vectors (3 * nRows)
    int ndofFree = 0;                   // free DOFs (the reference's numColumns)
    int nRows = 0;                      // block rows
    bool partial = false;               // some node has one or two constrained DOFs
    std::vector<int> nodeRow;           // node -> block row of its first free DOF, or -1
    std::vector<int> rowNode;           // block row -> node of its first DOF
    std::vector<int> nodeDof;           // partial only: 3 per node, free DOF index or -1

    // ---- canonical pattern (core/sparseMatrix.h:186-274 layout) -------------
    std::vector<int> rowPtr;            // blockPcol: size nRows+1
    std::vector<int> colIdx;            // block column of each stored block (blockRows/3)
    std::vector<int> valOff;            // blockColumnOris: size nRows+1
    std::vector<int> blkOff;            // first value index of each block
    int64_t nBlocks = 0, nValues = 0;

    // ---- transposed index: for row I, blocks (K, I) with K < I -------------
    std::vector<int> tPtr, tRow, tBlk;
    std::vector<int> tPos;              // per block: index in the transposed list, -1 for diagonal blocks

    // ---- element maps ---------------------------------------------------------
    std::vector<int> elemMap78;         // 78 per tet (fullOrderNonlinear.h:73-110)
    std::vector<int> pairBlk;           // 10 per tet: destination block or -1
    std::vector<uint8_t> pairT;         // 10 per tet: 1 when the element block lands transposed
    std::vector<int> pairSlot;          // 10 per tet: contribution slot (destination-sorted) or -1
    std::vector<int> blkCStart;         // size nBlocks+1: contribution range of each block
    std::vector<int> nodeSlot;          // 4 per tet: node-force contribution slot or -1
    std::vector<int> rowCStart;         // size nRows+1: node-force contribution range per row
    int64_t nPairContrib = 0, nNodeContrib = 0;

    // ---- partial only: value-level gathers and the reference's own layout -------
    // The reference stores the last DOF triple as truncated blocks and scalar "remained"
    // columns (core/sparseMatrix.h:90-274); the kernels keep full 3x3 blocks.
    std::vector<int> vPtr, vSrc;        // per value: contribution entries, pair slot * 9 + entry (-1: padding diagonal, 1.0)
    std::vector<int> dPtr, dSrc;        // per DOF: node-force entries, node slot * 3 + axis
    std::vector<int> refPcol, refOris, refRows;   // the reference's blockPcol, blockColumnOris, blockRows
    std::vector<int> refPos;            // per value: the reference's value index, or -1 (padding)
    int64_t nValuesRef = 0;             // values of the reference's layout (nValues when not partial)

    // ---- IC(0) level schedule -------------------------------------------------
    std::vector<int> level;             // per row
    std::vector<int> levelPtr;          // size nLevels+1
    std::vector<int> levelRows;         // rows sorted by level (stable)
    bool levelsContiguous = false;      // each level is the consecutive rows [levelPtr[l], levelPtr[l+1])
    int nLevels = 0;

    // ---- instances and row chunks ---------------------------------------------
    int nInst = 1;
    std::vector<int> instRowStart;      // size nInst+1
    std::vector<int> instChunkStart;    // size nInst+1
    std::vector<int> chunkRow0, chunkRow1, chunkInst;
    std::vector<int> chunkSorted;       // chunks by (instance, centroid along the longest axis)
    int perRun = 0;                     // chunkSorted is cut into nXcd runs of perRun chunks (last shorter)
    std::vector<int> rowInst;           // per row
    std::vector<int> tetInst;           // per tet
};

// Builds every derived structure.  Returns an empty string or an error text.
std::string buildTopology(Topology &t, int chunkRows, int nXcd = 8);

// Layout of the grouped symmetric product (od_groups.cpp).  Up to kGroupChunks chunks of one
// instance that share many blocks form a group, one workgroup of the product with one wave
// per chunk (local row = wave * chunkRows + row in chunk).  The group stores each upper block
// of its rows once; a block whose column row is in the group is "internal": its product for
// the column row (the transposed block times the row's x) goes to an LDS deposit, so the
// block is read once for both rows.  The lower slots of the group's rows whose upper owner
// lies outside the group come from the group's crossing list.  Every row then sums its slots
// (diagonal, upper blocks, lower blocks in transposed-list order) exactly as the other SpMV
// forms do.  Rows with more than 16 blocks are not supported (ok = false).
constexpr int kGroupChunks = 4;
struct GroupLayout {
    bool ok = false;
    std::string why;                 // reason when !ok
    int nGroups = 0, perRun = 0;     // groups in run order, cut into nXcd runs of perRun
    int maxDeps = 0;                 // most LDS deposits (lower slots) of any group
    int maxHalo = 0;                 // most halo rows of any group
    int64_t nUp = 0, nX = 0;         // upper-stream positions (padded per chunk), crossing positions (padded to 64)
    std::vector<int> runChunks;      // 4 per run position: the group's chunks (-1: none)
    std::vector<int> runX;           // 4 per run position: crossing base, crossing count, instance, 0
    std::vector<int> chunkUp;        // 2 per chunk: upper-stream base, width (slots)
    std::vector<int> rowInfo;        // per row: nUp | items << 4 | depBase << 9
    std::vector<int> upMeta;         // per upper position: -1 padding; >= 0 crossing (the column
                                     // row's slot in the group's halo); else 0x80000000 | local
                                     // column << 14 | deposit
    std::vector<int> upSrc;          // per upper position: value offset of the block (-1: padding)
    std::vector<int> xMeta;          // 2 per crossing position: halo slot of the column row (the upper
                                     // owner), deposit (-1: padding)
    std::vector<int> grpHalo;        // 2 per run position: base in haloRows, count
    std::vector<int> haloRows;       // per group, in row order: the rows outside the group whose x
                                     // its upper blocks and crossing blocks multiply
    std::vector<int> xSrc;           // per crossing position: value offset of the block (-1: padding)
};
std::string buildGroupLayout(const Topology &t, int chunkRows, int nXcd, GroupLayout &g);

// GPS relabelling (od_relabel.cpp): new label per node, elements relabelled and sorted by
// label sum in place; returns an error message or nullptr.
const char *relabelGps(int nNodes, int nElems, int npe, int *elements, int *labels);

// Global DOF of (node, axis): -1 when constrained (nodeIndexer.cpp:12-25).
int globalIndex(const Topology &t, int node, int axis);

}  // namespace od
