// Internal structures shared by the kernels and the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "od_precond.h"
#include "od_setup.h"

namespace od {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define OD_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::od::HipError(std::string(#call) + " failed: " + hipGetErrorString(e_));            \
    } while (0)

constexpr int kChunkRows = 64;      // block rows per workgroup chunk (never straddles an instance)
constexpr int kBlock = 256;         // threads per workgroup
constexpr int kValuePad = 16;       // trailing entries allocated after every value array
// SpMV slots: 16 lanes per block row; slot k of row r holds the value offset and the column
// of the row's k-th block (diagonal first, then the upper blocks, then the transposed ones)
constexpr int kSpmvLanes = 16;
constexpr int kSpmvThreads = kSpmvLanes * kChunkRows;
constexpr int kSlotDiag = 1 << 29, kSlotTrans = 1 << 30, kSlotColMask = (1 << 29) - 1;
constexpr int kRunLong = 1 << 30, kRunInstMask = kRunLong - 1;   // run record: a row of the chunk has more than kSpmvLanes blocks

// Device copy of the topology and material.
struct DevMesh {
    int nNodes, nTets, nRows, nInst, nChunks, nLevels;
    int64_t nBlocks, nValues, nT;   // nT: entries of the transposed index (lower blocks)
    // FP32 sweep slots: per chunk, the lower (forward) and the upper (backward) off-diagonal
    // blocks of its rows slot-major (position k * 64 + r: block k of the chunk's row r), each
    // chunk as wide as its longest row; swNF / swNB positions in all
    int64_t swNF, swNB;
    const int4 *swChunk;  // per chunk: {forward offset, forward width, backward offset, backward width}
    int *swColF, *swColB; // per position: the column's slot in the chunk's halo (-1: padding)
    // per chunk and direction its halo: the rows whose x its swept-side blocks multiply, in
    // row order, at a fixed stride (swHF / swHB rows per chunk, padded with row 0), so that a
    // sweep workgroup loads them with the chunk's record and stages their x in LDS
    int *swHaloF, *swHaloB;
    int swHF, swHB;
    const int2 *swHaloCnt;   // per chunk: the halo's rows forward, backward
    int *swBaseF, *swBaseB;   // per row: its position of block 0 (offset + r)
    int *levelChunks;     // per level, the chunks holding its rows, in chunk order
    const int *levelChunkPtr;   // host memory: level l's chunks are levelChunks[ptr[l], ptr[l + 1])
    double *X;            // 3*nNodes
    int *tets;            // 4*nTets
    int *nodeRow;         // nNodes
    // partially constrained nodes (else NULL): free DOF of each node axis (-1: constrained),
    // and the value-level / DOF-level gather lists that replace the block-level ones
    int *nodeDof;         // 3*nNodes
    int *vPtr, *vSrc;     // nValues+1 / entries: pair slot * 9 + entry, -1 for a padding diagonal (1.0)
    int *dPtr, *dSrc;     // 3*nRows+1 / entries: node slot * 3 + axis
    int *rowPtr, *colIdx, *blkOff, *blkRow;     // pattern
    int *valOff;                                 // first value of each block row (nRows+1)
    int *tPtr, *tRow, *tBlk, *tOff;             // transposed index (+ value offset of each block)
    int *tPos;                                   // per block: its place in the transposed list (-1: diagonal)
    int *pairBlk, *pairSlot; uint8_t *pairT;    // 10 per tet
    // corotational tets in compact form (no partially constrained nodes): per contribution slot
    // its source (tet << kSrcTetShift | stored flag | pair << 1 | transposed; the flag marks a
    // tet that is not corotational, whose block the element pass stores); per tet, from the
    // last stiffness pass, the
    // rotated shape-function gradients g_q = R dN_q (12) and lambda / (36 V0), mu / (36 V0).
    // The element pass stores these instead of the 10 blocks, and the block gather forms the
    // blocks R K0_ab R^T from them.
    int *slotSrc;
    double *rotC;                                // 16 per tet (NULL: no compact form)
    uint8_t *blkCompact;                         // per block: has a corotational contribution (with rotC)
    // the compact gather's work: 8 runs (one per XCD) of gatherPerRun workgroups of kBlock
    // segments each, the blocks of the SpMV runs' chunks in order, each block cut into segments
    // of at most kGatherSeg contributions that one workgroup holds: {block, first and end
    // contribution, segments that follow (first segment of the block) or -1}; padding block -1
    const int4 *gatherSeg;
    int gatherPerRun;
    int *blkCStart;                              // nBlocks+1
    int *nodeSlot, *rowCStart;                   // 4 per tet / nRows+1
    int *levelRows;                              // nRows
    int levelsContiguous;                        // levels are consecutive row ranges (levelRows is the identity)
    int sweepDotLists;                           // several instances, short level chunk lists: the sweeps
                                                 // that finish z form the r.z chunk partials
    const int *chunkTopLevel;                    // per chunk, the highest level of its rows
    int *rowLevel;                               // IC(0) dependency level of each row
    // IC(0) update pairs of each upper block b = (I, J): the lower neighbours K of I (in
    // transposed-list order) whose row also holds column J, as value offsets of U_KI and U_KJ
    int *icPairPtr;                              // nBlocks+1
    const int2 *icPairs;
    // per row, for the sweeps: {tPtr[r], tPtr[r+1], valOff[r], instance} and
    // {rowPtr[r], rowPtr[r+1], valOff[r], instance}
    const int4 *rowFwd, *rowBwd;
    // per level, in host memory (read at launch, never on the device): the form of the forward
    // and backward sweeps (lanes per row 4, 8 or 16 | slot kernel selector << 5), for the FP32
    // copies and for the FP64 copies
    const uint8_t *levelLanesFwd, *levelLanesBwd, *levelLanesFwd64, *levelLanesBwd64;
    int *chunkRow0, *chunkRow1, *chunkInst;      // nChunks
    // SpMV run order: the chunks in spatial order, cut into 8 runs of perRun positions (one
    // run per XCD); per position {chunk, first row, end row, instance | kRunLong}
    const int4 *runInfo;
    int perRun;
    int spmvR64, spmvR32; // chunks per SpMV workgroup (FP64 / FP32 values), set by the autotuner
    int spmvL64, spmvL32; // lanes per block row of the SpMV (16, 8 or 4), set by the autotuner
    int *slotOff, *slotCol;                      // kSpmvThreads per run position (padding: slotCol = -1)
    int *svCol;           // slotCol slot-major (position s * 64 + r: slot s of row r), the slot-value order
    // best shape among the chunk-descriptor forms (the FP64 product when no grouped values
    // are at hand, e.g. y = A x outside a solve)
    int spmvR64d, spmvL64d;
    // grouped symmetric product (GroupLayout in od_setup.h; lanes -1 selects it): per run
    // position the group's chunks and {crossing base, crossing count, instance, 0}; per chunk
    // {upper-stream base, width}; per row nUp | items << 4 | depBase << 9; per upper position
    // its meta; per crossing position {column row, deposit}; the fill sources of both
    int grpOk64, grpOk32;   // the form fits (rows of at most 16 blocks, LDS within 64 KB)
    int nGroups, perRunG, grpMaxDeps, grpMaxHalo;
    int64_t grpNu, grpNx;
    const int4 *grpChunks, *grpX;
    const int2 *grpChunkUp, *grpXMeta, *grpHalo;
    const int *grpRowInfo, *grpUpMeta, *grpUpSrc, *grpXSrc, *grpHaloRows;
    int *instChunkStart, *instRowStart;          // nInst+1
    int *tetInst, *rowInst;
    // material (per tet)
    int *kind; double *E, *nu, *rho;
    uint8_t hasKind[3];   // host-side: some tet is corotational / neo-Hookean / corotational plastic
    // corotational plastic: per-tet yield criterion, creep ratio, maximum plastic strain, and the
    // plastic strain state (9 per tet), updated by every stiffness and force evaluation
    double *plYield, *plCreep, *plMax, *plStrain;
    double thrCorot, thrNeo;
};

// Per-instance PCG scalars (device resident).
struct InstScalars {
    double delta, bb, rr, eps, alpha, beta, pq, rz;
    int it, done, maxIt, pad;
};

// What the element kernel writes per node (see ElemParams).
struct ElemParams {
    int mode;             // 0: stiffness blocks + node terms, 1: mass blocks only
    int nodeTerms;        // write node contributions
    double cyv, cya;      // y = cyv*v + cya*a  (multiplied by K_e)
    double czv, cza, cm;  // z = czv*v + cza*a  (multiplied by cm*M_e)
};

// contribution source encoding of the compact gather (DevMesh::slotSrc)
constexpr int kSrcTetShift = 6, kSrcStored = 32;
// contributions per segment of the compact gather (a diagonal block has about 24)
constexpr int kGatherSeg = 4;

// ---- kernel launchers (od_kernels.hip) ------------------------------------------
void launch_element(const DevMesh &m, hipStream_t s, const double *d, const double *v, const double *a,
                    const ElemParams &p, double *blkC, double *nodeC, const uint8_t *instActive);
// sets the stored flag of each of the nContrib contribution sources from the tets' kinds
void launch_mark_stored(const DevMesh &m, hipStream_t s, int64_t nContrib);
// stiff: the contributions come from a stiffness pass (corotational tets in compact form are
// formed from their rotations); otherwise (mass pass) every contribution is a stored block
void launch_gather_blocks(const DevMesh &m, hipStream_t s, const double *blkC, double cK, const double *M, double cM,
                          double *out, int accumulate, float *out32, const uint8_t *instActive, bool stiff);
void launch_gather_rhs(const DevMesh &m, hipStream_t s, const double *nodeC, const double *fext, const double *imp,
                       double cf, double *out, int mode, const uint8_t *instActive);
void launch_body_force(const DevMesh &m, hipStream_t s, const double *bodyPerTet, double *nodeC, double *f);
void launch_expand_map(const DevMesh &m, hipStream_t s, int *map78);
// y = A x or y += A x.  FP64 y = A x runs the autotuned FP64 shape: the grouped form when it
// is selected and its values (grpVals) are given, else the best chunk-descriptor shape.
template <typename T>
void launch_spmv(const DevMesh &m, hipStream_t s, const T *vals, const T *x, T *y, int accumulate,
                 const InstScalars *sc, const T *grpVals = nullptr);
// PCG SpMV with R chunks per workgroup (1, 2 or 4), timed by optional dispatch events
// With z set (the PCG product), x is the previous direction and the product runs on the new
// direction p = z + beta x of each instance, written to pNew (see spmv_chunks).
// LS = 0 selects the FP32 slot-value form (slotVals, filled by launch_fill_slot_values from
// vals; R is then the slots per thread, 1, 2 or 4); vals still serves rows longer than
// kSpmvLanes blocks.  LS = -1 selects the grouped form (FP64 or FP32; slotVals then holds the
// grouped values of launch_fill_grp; R is the slots per thread, 4 or 8).
template <typename T>
void launch_spmv_dot(const DevMesh &m, hipStream_t s, int R, int LS, const T *vals, const T *x, T *y, double *partials,
                     const InstScalars *sc, hipEvent_t start = nullptr, hipEvent_t stop = nullptr,
                     const T *z = nullptr, T *pNew = nullptr, const T *slotVals = nullptr);
constexpr int kSpmvVariants[3] = {1, 2, 4};        // chunks per workgroup (slot-value form: slots per thread)
constexpr int kSpmvLaneVariants[3] = {16, 8, 4};   // lanes per block row
// FP32 slot values: per run position, the kSpmvThreads slot blocks as the product multiplies
// them (diagonal blocks full, lower blocks transposed, padding zero), slot-major (position
// s * 64 + r: slot s of row r) and part-major: entries 0-3 of every slot (float4 each), then
// entries 4-7, then entry 8
constexpr int kSlotValFloats = 9 * kSpmvThreads;
void launch_fill_slot_values(const DevMesh &m, hipStream_t s, const float *vals, float *slotVals);
// Grouped product values (lanes -1): the upper stream (9 entries per position), the crossing
// list (9 per position) and the packed diagonals (6 per position, 64 positions per chunk),
// each part-major (entries in vector planes: FP64 pairs, FP32 quads, then the rest), refreshed
// from the canonical values at every solve that runs the form.
constexpr int kGrpThreads = kGroupChunks * kChunkRows;   // rows of a group (threads per 16 slots per thread)
constexpr int kGrpSlotVariants[2] = {4, 8};               // slots per thread of the grouped product
inline int64_t grp_values(const DevMesh &m) {
    return 9 * (m.grpNu + m.grpNx) + 6 * (int64_t)m.nChunks * kChunkRows;
}
// LDS a workgroup may hold on the current device (the device's per-block limit; 64 KB if the
// query fails)
size_t lds_max();
// LDS of the grouped product with K slots per thread: the group's x, its halo's x, then the
// deposits, whose area afterwards holds the cross-thread row sums (16 / K - 1 per row)
template <typename T>
size_t grp_lds_bytes(const DevMesh &m, int K) {
    const size_t red = (size_t)(16 / K - 1) * kGrpThreads;
    return sizeof(T) * 3 * ((size_t)kGrpThreads + (size_t)m.grpMaxHalo + std::max((size_t)m.grpMaxDeps, red));
}
template <typename T>
void launch_fill_grp(const DevMesh &m, hipStream_t s, const T *vals, T *grpVals);
// IC(0) factor copies of the sweeps (FP32 for the mixed form, FP64 for the FP64 sweeps),
// part-major: FP32 position b of an array of nb positions has entries 0-3 at [4 b, 4 b + 4),
// entries 4-7 at [4 nb + 4 b, ...) and entry 8 at [8 nb + b]; FP64 has the pairs (0, 1), (2,
// 3), (4, 5), (6, 7) at [2 q nb + 2 b, ...) and entry 8 at [8 nb + b]; so a wave's loads of
// consecutive positions read consecutive memory.  The upper copy holds the diagonal blocks
// first (8 entries per row: the 6 packed entries, 2 unused), then the upper blocks at their
// backward sweep slots (swNB positions); the lower copy holds the lower blocks at their
// forward sweep slots (swNF positions).  Padding positions stay zero.  Sizes in entries:
inline int64_t ic0_u32_floats(const DevMesh &m) { return 8 * (int64_t)m.nRows + 9 * m.swNB; }
inline int64_t ic0_ul32_floats(const DevMesh &m) { return 9 * m.swNF; }

// PCG building blocks (see od_kernels.hip for the exact phase semantics)
struct PcgWork {
    double *partials;     // 3 * nChunks, slot q of chunk c at [q * nChunks + c] (slots: r.z or p.q, r.r, b.b)
    InstScalars *sc;      // nInst
    int *doneFlags;       // nInst, mirrors sc[i].done for host polling
    // inner solves of the mixed form: the outer scalars, whose eps (tol^2 |b|^2 of the
    // FP64 system) bounds the inner target from below (see k_finalize, PH_INIT)
    const InstScalars *outer = nullptr;
    const void *slotVals = nullptr;   // FP32 slot values of the matrix (slot-value product form)
};
template <typename T>
void pcg_init_residual(const DevMesh &m, hipStream_t s, const T *b, const T *q, T *r, PcgWork &w);
template <typename T>
void pcg_jacobi(const DevMesh &m, hipStream_t s, const T *invDiag, const T *r, T *z, PcgWork &w);
template <typename T>
void pcg_dot_rz_init(const DevMesh &m, hipStream_t s, const T *r, const T *z, T *p, PcgWork &w, double tol, int maxIt,
                     const uint8_t *preDone);
// p = z + beta p (the direction update when the product does not fuse it)
template <typename T>
void pcg_update_p(const DevMesh &m, hipStream_t s, const T *z, T *p, PcgWork &w);
// q = A p; with z set, p = z + beta pOld is computed on the fly and written to p (the
// direction update fused into the product), else the product runs on p = pOld
template <typename T>
void pcg_alpha(const DevMesh &m, hipStream_t s, const T *vals, const T *pOld, const T *z, T *p, T *q, PcgWork &w,
               hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
void pcg_alpha_finalize(const DevMesh &m, hipStream_t s, PcgWork &w);
// The IC(0) iteration forms the r.z chunk partials inside its sweeps (one instance, levels as
// consecutive row ranges, at least two levels): the iteration then ends with
// pcg_finalize_iter instead of pcg_dot_rz.
bool ic0_dot_fused(const DevMesh &m, const std::vector<int> &levelPtr);
void pcg_finalize_iter(const DevMesh &m, hipStream_t s, PcgWork &w);
template <typename T>
void pcg_update_xr(const DevMesh &m, hipStream_t s, const T *p, const T *q, T *x, T *r, const T *invDiag, T *z,
                   PcgWork &w);
template <typename T>
void pcg_dot_rz(const DevMesh &m, hipStream_t s, const T *r, const T *z, PcgWork &w);
// IC(0)-preconditioned iteration: x += alpha p, r -= alpha q with the forward solve of the
// level-0 rows fused in, then the remaining sweep levels (the backward solve of the last
// level fused into its forward launch).  Same result as pcg_update_xr + launch_ic0_apply.
template <typename T>
void pcg_update_xr_ic0(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, const T *p, const T *q, T *x,
                       T *r, const T *U, const T *Ul, T *z, PcgWork &w, const uint8_t *skip);
template <typename T>
void launch_jacobi_diag(const DevMesh &m, hipStream_t s, const double *vals, T *invDiag);
// sweep form of each level for factor copies of type T (FP32: the mixed form's sweeps, FP64:
// the FP64 sweeps); codes as launch_sweep reads them
template <typename T>
void tune_ic0_lanes(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, uint8_t *fwd, uint8_t *bwd);
// one read of a large buffer (nFloats floats) that evicts the caches before a timed launch
void launch_stream_read(hipStream_t s, float *buf, int64_t nFloats);
// IC(0) factor U of A (canonical layout) and, when given, its part-major sweep copies of type T
// (Uc: diagonal blocks, 8 entries per row, then the upper blocks at their backward sweep
// slots; Ulc: the lower blocks at their forward sweep slots; sizes ic0_u32_floats /
// ic0_ul32_floats entries), which the sweeps read
template <typename T>
void launch_ic0_factor(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, const double *A, double *U,
                       T *Uc, T *Ulc, double pivotEps, const uint8_t *skip);
template <typename T>
void launch_ic0_apply(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, const T *U, const T *Ul,
                      const T *r, T *z, const InstScalars *sc, const uint8_t *skip);
void launch_cast_f64_f32(hipStream_t s, const double *a, float *b, int64_t n);
void launch_residual_f64(const DevMesh &m, hipStream_t s, const double *b, const double *Ax, double *r, float *r32,
                         PcgWork &w, double tol, const uint8_t *preDone, int first);
void launch_add_correction(const DevMesh &m, hipStream_t s, const float *dx, double *x, const InstScalars *outer);
void launch_newmark_correct(const DevMesh &m, hipStream_t s, double *d, double *v, const double *aOld,
                            const double *aNew, double h, double beta, double gamma, const uint8_t *instActive);
// The step graph's corrector: instances whose flag in `commit` is set get the Newmark
// correction and a = aNew; the others keep their state (the eager step then completes them).
void launch_newmark_commit(const DevMesh &m, hipStream_t s, double *d, double *v, double *a, const double *aNew,
                           double h, double beta, double gamma, const uint8_t *commit);
void launch_be_correct(const DevMesh &m, hipStream_t s, double *v, const double *dv, const uint8_t *instActive);
void launch_axpy_rows(const DevMesh &m, hipStream_t s, double *y, double a, const double *x, const uint8_t *instActive);

// ---- plane collision damping (od_damping.hip) --------------------------------------
// Per-call buffers sized for nPairs, and per-block counters that are zero between calls.
struct DampScratch {
    int *cnt = nullptr, *fill = nullptr, *start = nullptr;   // nBlocks each
    double *planes = nullptr;                                  // 4 per pair (normal, constant)
    int *tets = nullptr;                                       // 1 per pair
    double *pv = nullptr;                                      // 90 per pair
    int *dst = nullptr, *list = nullptr, *touched = nullptr;   // 10 per pair each
    int *nTouched = nullptr;
};
// D += the damping matrix of nPairs (plane, tet) pairs (w.planes, w.tets already on the
// device); disp and vel are free-DOF vectors (NULL: zero)
void launch_plane_damping(const DevMesh &m, hipStream_t s, int nPairs, const DampScratch &w, double coeff,
                          const double *disp, const double *vel, double *D);
void launch_set_done_from_active(const DevMesh &m, hipStream_t s, uint8_t *preDone, const uint8_t *instActive);

// ---- scalar preconditioner sweeps (od_precond.hip) ----------------------------------
// Device copy of a ScalarSweepsHost (arrays owned by the solver).
struct ScalarSweepsDev {
    int n = 0;
    bool unit = false;
    const int *fPtr = nullptr, *fCol = nullptr, *bPtr = nullptr, *bRow = nullptr;
    const double *fVal = nullptr, *bVal = nullptr, *diag = nullptr;
    const int *fLevelPtr = nullptr, *fLevelRows = nullptr, *bLevelPtr = nullptr, *bLevelRows = nullptr;
    const ScalarSweepPlan *fPlan = nullptr, *bPlan = nullptr;   // host memory
};
// z = (L D L^T)^-1 r or (L L^T)^-1 r; z may not alias r.
void launch_scalar_apply(const ScalarSweepsDev &s, hipStream_t st, const double *r, double *z);

}  // namespace od
