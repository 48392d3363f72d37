// Plane collision damping: the damping matrix of MeshPlaneCollisionDampingForcer
// (collision/penaltyDampingForcer.cpp:8-61, collision/intersection.cpp:12-143) for a list
// of (plane, tet) pairs, added into a matrix in the canonical layout.
//
// The reference adds the pairs one after another with +=, so an entry that several pairs
// touch sums their contributions in pair order.  Here one thread evaluates one pair (its
// 10 blocks, already in destination orientation), a counting sort groups the block
// contributions by destination, and one thread per touched block adds its contributions
// sorted by (pair, block) onto the stored value: the same order as the reference, and the
// same bits on every run.  The work is proportional to the number of pairs, not the mesh.
//
// Floating-point contraction is off in this file (the pragma precedes the includes, so it
// covers the vector helpers too): every product and sum rounds as in the reference's scalar code.
#pragma clang fp contract(off)
#include "od_elem.h"
#include "od_internal.h"

namespace od {

namespace {

using namespace odk;

constexpr int kDampThreads = 128;

// intersection.cpp:12-17: where segment pa-pb crosses the plane, t clamped to [0, 1]
__device__ v3 seg_plane(v3 n, double c, v3 pa, v3 pb) {
    const v3 ab = pb - pa;
    double t = (-c - dot(n, pa)) / dot(n, ab);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return pa + ab * t;
}

__device__ v3 unit(v3 v) {   // latool.h:528-530: v / |v|
    return vdiv(v, sqrt(dot(v, v)));
}

__device__ v3 neg(v3 v) { return mk(-v.x, -v.y, -v.z); }

// intersection.cpp:20-117: centroid of the part of the tet below the plane and dir, six
// times that part's volume along the normal (points within 1e-6 above count as below)
__device__ void plane_tet(v3 n, double c, const v3 *p, v3 &centroid, v3 &dir) {
    const double eps = 1e-6;
    v3 ng[4], ps[4];
    int nn = 0, np = 0;
    for (int i = 0; i < 4; i++) {
        if (dot(n, p[i]) + c < eps) ng[nn++] = p[i];
        else ps[np++] = p[i];
    }
    if (nn == 0) {
        double minDis = 1.7976931348623157e308;
        for (int i = 0; i < 4; i++) {
            const double dis = dot(n, p[i]) + c;
            if (dis < minDis) { minDis = dis; centroid = p[i]; }
        }
        dir = mk(0, 0, 0);
    } else if (nn == 1) {
        const v3 q0 = seg_plane(n, c, ng[0], ps[0]), q1 = seg_plane(n, c, ng[0], ps[1]), q2 = seg_plane(n, c, ng[0], ps[2]);
        centroid = (ng[0] + q0 + q1 + q2) * 0.25;
        const v3 ab = q0 - ng[0], ac = q1 - ng[0], ad = q2 - ng[0];
        const double svol = dot(ab, cross(ac, ad));
        dir = unit(cross(ab, ac) + cross(ac, ad) + cross(ad, ab)) * fabs(svol);
        if (svol < 0.0) dir = neg(dir);
    } else if (nn == 2) {
        const v3 q0 = seg_plane(n, c, ng[0], ps[0]), q1 = seg_plane(n, c, ng[0], ps[1]);
        const v3 q2 = seg_plane(n, c, ng[1], ps[0]), q3 = seg_plane(n, c, ng[1], ps[1]);
        const v3 ab = ng[1] - ng[0];
        // the wedge below the plane as the tets (n0 n1 q2 q3), (n0 q0 q1 q2), (n0 q1 q2 q3)
        const v3 c0 = ng[0] + ng[1] + q2 + q3;
        const double v0 = fabs(dot(ng[1] - ng[0], cross(q2 - ng[0], q3 - ng[0])));
        const v3 c1 = ng[0] + q0 + q1 + q2;
        const double v1 = fabs(dot(q0 - ng[0], cross(q1 - ng[0], q2 - ng[0])));
        const v3 c2 = ng[0] + q1 + q2 + q3;
        const double v2 = fabs(dot(q1 - ng[0], cross(q2 - ng[0], q3 - ng[0])));
        centroid = vdiv(c0 * v0 + c1 * v1 + c2 * v2, v0 + v1 + v2) * 0.25;
        v3 d = cross(q0 - ng[0], q1 - ng[0]) + cross(q3 - ng[1], q2 - ng[1]) + cross(ab, q2 - ng[0]) +
               cross(q2 - ng[0], q0 - ng[0]) + cross(q3 - ng[0], ab) + cross(q1 - ng[0], q3 - ng[0]);
        dir = unit(d) * (v0 + v1 + v2);
        if (dot(ng[1] - ng[0], cross(ps[0] - ng[0], ps[1] - ng[0])) < 0.0) dir = neg(dir);
    } else if (nn == 3) {
        const v3 q0 = seg_plane(n, c, ng[0], ps[0]), q1 = seg_plane(n, c, ng[1], ps[0]), q2 = seg_plane(n, c, ng[2], ps[0]);
        const v3 ab = q0 - ps[0], ac = q1 - ps[0], ad = q2 - ps[0];
        const v3 cs = p[0] + p[1] + p[2] + p[3];
        const double sv = dot(ng[0] - ps[0], cross(ng[1] - ps[0], ng[2] - ps[0]));
        const double v = fabs(sv);
        const v3 c0 = ps[0] + q0 + q1 + q2;
        const double v0 = fabs(dot(ab, cross(ac, ad)));
        // the tet minus the small tet above the plane at ps[0]
        centroid = vdiv(cs * v - c0 * v0, v - v0) * 0.25;
        dir = unit(cross(ab, ac) + cross(ac, ad) + cross(ad, ab)) * (v - v0);
        if (sv > 0.0) dir = neg(dir);
    } else {
        centroid = (p[0] + p[1] + p[2] + p[3]) * 0.25;
        const double vol = fabs(dot(p[1] - p[0], cross(p[2] - p[0], p[3] - p[0])));
        dir = n * vol;
    }
}

// intersection.cpp:119-143: unsigned sub-volume ratios; 1/4 each when 6V <= 1e-6
__device__ void barycentric(const v3 *p, v3 q, double *co) {
    const v3 ab = p[1] - p[0], ac = p[2] - p[0], ad = p[3] - p[0], bc = p[2] - p[1], bd = p[3] - p[1];
    const v3 ap = q - p[0], bp = q - p[1];
    const double va = fabs(dot(bp, cross(bc, bd))), vb = fabs(dot(ap, cross(ac, ad))), vc = fabs(dot(ap, cross(ab, ad))),
                 vd = fabs(dot(ap, cross(ab, ac)));
    const double vol = fabs(dot(ab, cross(ac, ad)));
    if (vol <= 1e-6) { co[0] = co[1] = co[2] = co[3] = 0.25; return; }
    const double inv = 1.0 / vol;
    co[0] = va * inv; co[1] = vb * inv; co[2] = vc * inv; co[3] = vd * inv;
}

// One thread per pair: the pair's 10 node-pair blocks (9 values each; a diagonal block uses
// the first 6, in the packed order) in destination orientation, their destination blocks
// (-1: constrained node or no damping), and the per-block contribution counts.
__global__ __launch_bounds__(kDampThreads) void k_damp_eval(DevMesh m, int nPairs, const double *__restrict__ planes,
                                                            const int *__restrict__ ptet, double coeff,
                                                            const double *__restrict__ disp,
                                                            const double *__restrict__ vel, double *__restrict__ pv,
                                                            int *__restrict__ dst, int *__restrict__ cnt,
                                                            int *__restrict__ touched, int *__restrict__ nTouched) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nPairs) return;
    const int e = ptet[k];
    const v3 n = mk(planes[4 * k], planes[4 * k + 1], planes[4 * k + 2]);
    const double c = planes[4 * k + 3];
    v3 p[4], w[4];
    for (int a = 0; a < 4; a++) {
        const int nd = m.tets[4 * e + a], r = m.nodeRow[nd];
        v3 u = mk(0, 0, 0);
        w[a] = mk(0, 0, 0);
        if (r >= 0) {
            if (disp) u = mk(disp[3 * r], disp[3 * r + 1], disp[3 * r + 2]);
            if (vel) w[a] = mk(vel[3 * r], vel[3 * r + 1], vel[3 * r + 2]);
        }
        p[a] = mk(m.X[3 * nd], m.X[3 * nd + 1], m.X[3 * nd + 2]) + u;
    }
    v3 centroid, dir;
    plane_tet(n, c, p, centroid, dir);
    double co[4];
    barycentric(p, centroid, co);
    v3 nd = mk(0, 0, 0);
    if (dot(dir, dir) != 0.0) nd = unit(dir);
    v3 vc = mk(0, 0, 0);
    for (int a = 0; a < 4; a++) vc = vc + w[a] * co[a];
    const bool on = dot(vc, nd) < 0.0;
    const v3 cd = dir * coeff;
    const double t[9] = {cd.x * nd.x, cd.x * nd.y, cd.x * nd.z, cd.y * nd.x, cd.y * nd.y,
                         cd.y * nd.z, cd.z * nd.x, cd.z * nd.y, cd.z * nd.z};
    int q = 0;
    for (int a = 0; a < 4; a++)
        for (int b = a; b < 4; b++, q++) {
            const int slot = 10 * k + q;
            const int blk = on ? m.pairBlk[10 * e + q] : -1;
            dst[slot] = blk;
            if (blk < 0) continue;
            double *o = pv + (size_t)slot * 9;
            if (a == b) {
                // stored entries (0,0) (1,0) (2,0) (1,1) (2,1) (2,2) of the lower diagonal block
                // receive t(0,0) t(0,1) t(0,2) t(1,0) t(1,1) t(2,0) (penaltyDampingForcer.cpp:40-45)
                const double s = co[a] * co[a];
                o[0] = s * t[0]; o[1] = s * t[1]; o[2] = s * t[2];
                o[3] = s * t[3]; o[4] = s * t[4]; o[5] = s * t[6];
            } else {
                const double s = co[a] * co[b];
                const bool tr = m.pairT[10 * e + q] != 0;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) o[tr ? j * 3 + i : i * 3 + j] = s * t[i * 3 + j];
            }
            if (atomicAdd(&cnt[blk], 1) == 0) touched[atomicAdd(nTouched, 1)] = blk;
        }
}

// One workgroup: start offset of each touched block's bucket (exclusive prefix of the counts
// in touched-list order).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_damp_offsets(const int *__restrict__ touched,
                                                               const int *__restrict__ nTouched,
                                                               const int *__restrict__ cnt, int *__restrict__ start) {
    __shared__ int part[kScanThreads];
    const int n = *nTouched, per = (n + kScanThreads - 1) / kScanThreads;
    const int lo = min(n, (int)threadIdx.x * per), hi = min(n, lo + per);
    int s = 0;
    for (int i = lo; i < hi; i++) s += cnt[touched[i]];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
        const int v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int i = lo; i < hi; i++) {
        start[touched[i]] = run;
        run += cnt[touched[i]];
    }
}

// Places each contribution (key 10 * pair + block) into its destination's bucket.
__global__ __launch_bounds__(kDampThreads) void k_damp_fill(int nSlots, const int *__restrict__ dst,
                                                            const int *__restrict__ start, int *__restrict__ fill,
                                                            int *__restrict__ list) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nSlots) return;
    const int blk = dst[k];
    if (blk >= 0) list[start[blk] + atomicAdd(&fill[blk], 1)] = k;
}

// One thread per touched block: sorts its bucket by key and adds the contributions onto the
// stored values in that order; leaves the block's counters at zero for the next call.
__global__ __launch_bounds__(kDampThreads) void k_damp_apply(DevMesh m, int maxTouched, const int *__restrict__ touched,
                                                             const int *__restrict__ nTouched, int *__restrict__ cnt,
                                                             int *__restrict__ fill, const int *__restrict__ start,
                                                             int *__restrict__ list, const double *__restrict__ pv,
                                                             double *__restrict__ D) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= maxTouched || k >= *nTouched) return;
    const int blk = touched[k], n = cnt[blk];
    int *L = list + start[blk];
    for (int i = 1; i < n; i++) {
        const int key = L[i];
        int j = i - 1;
        while (j >= 0 && L[j] > key) { L[j + 1] = L[j]; j--; }
        L[j + 1] = key;
    }
    const int nq = m.colIdx[blk] == m.blkRow[blk] ? 6 : 9;
    double *out = D + m.blkOff[blk];
    for (int q = 0; q < nq; q++) {
        double acc = out[q];
        for (int i = 0; i < n; i++) acc += pv[(size_t)L[i] * 9 + q];
        out[q] = acc;
    }
    cnt[blk] = 0;
    fill[blk] = 0;
}

}  // namespace

void launch_plane_damping(const DevMesh &m, hipStream_t s, int nPairs, const DampScratch &w, double coeff,
                          const double *disp, const double *vel, double *D) {
    if (nPairs <= 0) return;
    const int nSlots = 10 * nPairs;
    OD_HIP(hipMemsetAsync(w.nTouched, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_damp_eval, dim3((nPairs + kDampThreads - 1) / kDampThreads), dim3(kDampThreads), 0, s, m,
                       nPairs, w.planes, w.tets, coeff, disp, vel, w.pv, w.dst, w.cnt, w.touched, w.nTouched);
    OD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_damp_offsets, dim3(1), dim3(kScanThreads), 0, s, w.touched, w.nTouched, w.cnt, w.start);
    OD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_damp_fill, dim3((nSlots + kDampThreads - 1) / kDampThreads), dim3(kDampThreads), 0, s, nSlots,
                       w.dst, w.start, w.fill, w.list);
    OD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_damp_apply, dim3((nSlots + kDampThreads - 1) / kDampThreads), dim3(kDampThreads), 0, s, m,
                       nSlots, w.touched, w.nTouched, w.cnt, w.fill, w.start, w.list, w.pv, D);
    OD_HIP(hipGetLastError());
}

}  // namespace od
