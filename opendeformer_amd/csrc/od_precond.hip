// Level-scheduled device sweeps of the scalar preconditioners (od_precond.h): one thread per
// row, each row a gather over its dependencies in the reference's order.  Floating-point
// contraction is off in this file (the pragma precedes the includes), so every product and
// difference rounds separately, as in the reference's scalar loops.
#pragma clang fp contract(off)

#include "od_internal.h"

namespace od {

template <bool UNIT>
__device__ __forceinline__ void forward_row(const ScalarSweepsDev &s, int i, const double *__restrict__ r,
                                            double *__restrict__ y) {
    double acc = r[i];
    for (int e = s.fPtr[i]; e < s.fPtr[i + 1]; e++) acc -= s.fVal[e] * y[s.fCol[e]];
    y[i] = UNIT ? acc : acc / s.diag[i];
}

template <bool UNIT>
__device__ __forceinline__ void backward_row(const ScalarSweepsDev &s, int i, double *__restrict__ z) {
    double dot = 0.0;
    for (int e = s.bPtr[i]; e < s.bPtr[i + 1]; e++) dot += z[s.bRow[e]] * s.bVal[e];
    z[i] = UNIT ? z[i] * s.diag[i] - dot : (z[i] - dot) / s.diag[i];
}

// One wide level: its rows spread over the grid.
template <bool FWD, bool UNIT>
__global__ __launch_bounds__(256) void k_scalar_level(ScalarSweepsDev s, int l, const double *__restrict__ r,
                                                      double *__restrict__ z) {
    const int *lp = FWD ? s.fLevelPtr : s.bLevelPtr, *lr = FWD ? s.fLevelRows : s.bLevelRows;
    const int q = lp[l] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= lp[l + 1]) return;
    if (FWD) forward_row<UNIT>(s, lr[q], r, z);
    else backward_row<UNIT>(s, lr[q], z);
}

// A run of narrow levels [l0, l1) in one workgroup: the rows of a level are spread over its
// threads, and a barrier (workgroup-scope release and acquire) separates the levels, so each
// level reads the finished rows of the earlier ones.
template <bool FWD, bool UNIT>
__global__ __launch_bounds__(1024) void k_scalar_levels(ScalarSweepsDev s, int l0, int l1, const double *__restrict__ r,
                                                        double *__restrict__ z) {
    const int *lp = FWD ? s.fLevelPtr : s.bLevelPtr, *lr = FWD ? s.fLevelRows : s.bLevelRows;
    for (int l = l0; l < l1; l++) {
        for (int q = lp[l] + (int)threadIdx.x; q < lp[l + 1]; q += (int)blockDim.x) {
            if (FWD) forward_row<UNIT>(s, lr[q], r, z);
            else backward_row<UNIT>(s, lr[q], z);
        }
        __syncthreads();
    }
}

template <bool FWD, bool UNIT>
static void sweep(const ScalarSweepsDev &s, const ScalarSweepPlan &p, hipStream_t st, const double *r, double *z) {
    for (const auto &L : p.launches) {
        if (L.wide) {
            const int cnt = p.levelPtr[L.l0 + 1] - p.levelPtr[L.l0];
            hipLaunchKernelGGL((k_scalar_level<FWD, UNIT>), dim3((cnt + 255) / 256), dim3(256), 0, st, s, L.l0, r, z);
        } else {
            hipLaunchKernelGGL((k_scalar_levels<FWD, UNIT>), dim3(1), dim3(1024), 0, st, s, L.l0, L.l1, r, z);
        }
        OD_HIP(hipGetLastError());
    }
}

void launch_scalar_apply(const ScalarSweepsDev &s, hipStream_t st, const double *r, double *z) {
    if (s.n == 0) return;
    // forward into z (y = L^-1 r), then the backward sweep in place on z
    if (s.unit) {
        sweep<true, true>(s, *s.fPlan, st, r, z);
        sweep<false, true>(s, *s.bPlan, st, nullptr, z);
    } else {
        sweep<true, false>(s, *s.fPlan, st, r, z);
        sweep<false, false>(s, *s.bPlan, st, nullptr, z);
    }
}

}  // namespace od
