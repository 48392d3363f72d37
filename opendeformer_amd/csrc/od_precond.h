// Scalar incomplete factorisations of the reference's preconditioners, built on the host from
// the canonical values (their fill depends on the values, so they are rebuilt at every solve
// as the reference's Preprocess does), and the schedule of their device sweeps (host code
// only; the device side is in od_internal.h and od_precond.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "od_setup.h"

namespace od {

// Lower factor in the reference's column storage: column c holds entries
// [pcol[c], pcol[c + 1]) of (rows, values).  Threshold IC stores the diagonal first in each
// column (solver/inCholeskyPreconditioner.cpp:40-64); SAINV LDL^T stores only the strictly
// lower entries of the unit factor, in the order it found them, and D^-1 separately
// (solver/inLDLTPreconditioner.cpp:88-101).  pcol has n + 1 entries.
struct ScalarFactor {
    int n = 0;
    std::vector<int> pcol, rows;
    std::vector<double> values;
    std::vector<double> invDiag;   // LDL^T only
};

// InCholeskyPreconditioner::incompleteCholeskyDecomposition (inCholeskyPreconditioner.cpp:83-130,
// with proccessSingleColumn :18-65): left-looking IC with drop tolerance eps and compensation
// relaxedScale, on the scalar lower triangle of A (whole-node constraints).
void thresholdIcFactor(const Topology &t, const double *vals, double eps, double relaxedScale, ScalarFactor &L);

// InLDLTPreconditioner::incompleteLDLTDecomposition (inLDLTPreconditioner.cpp:35-149): SAINV
// inverse-factor columns kept as row-linked node lists, entries above sainvEps kept in the
// inverse factor and above ldltEps in L.
void sainvLdltFactor(const Topology &t, const double *vals, double sainvEps, double ldltEps, ScalarFactor &L);

// Device form of a factor.  Forward: row i subtracts L(i, k) y_k for k < i in increasing k,
// then divides by L(i, i) (IC) or not (unit LDL^T).  Backward: row i forms the dot of its
// column's stored entries with the finished rows in stored order, then z_i = (y_i - dot) /
// L(i, i) (IC) or z_i = y_i D^-1_i - dot (LDL^T).  These are the operations, in the order, of
// the reference's solvePreconditionerSystem (inCholeskyPreconditioner.cpp:132-162,
// inLDLTPreconditioner.cpp:151-184), each row as a gather.  Rows are grouped by dependency
// level; a level of at least kScalarWideLevel rows is its own launch, and each run of
// narrower levels is one launch of a single workgroup that sweeps them with a barrier between
// levels.
constexpr int kScalarWideLevel = 1024;

struct ScalarSweepPlan {
    std::vector<int> levelPtr, levelRows;            // rows grouped by level
    struct Launch { int l0, l1; bool wide; };         // levels [l0, l1)
    std::vector<Launch> launches;
};

struct ScalarSweepsHost {
    int n = 0;
    bool unit = false;                               // LDL^T: unit diagonal, D^-1 in diag
    std::vector<int> fPtr, fCol, bPtr, bRow;
    std::vector<double> fVal, bVal, diag;
    ScalarSweepPlan fwd, bwd;
};
void buildScalarSweeps(const ScalarFactor &L, bool unit, ScalarSweepsHost &s);

}  // namespace od
