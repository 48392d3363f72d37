// HIP kernels for gfx950 (CDNA4, wave64) — the implicit-step hot path.
//
//  * element:   one thread per tet, FP64.  Writes the tet's 10 node-pair 3x3
//               blocks into a contribution array ordered by destination block,
//               and its 4 nodal 3-vectors into one ordered by node.
//  * gather:    one thread per stored value; sums the contributions of its block
//               in a fixed order (deterministic, atomic-free assembly) and fuses
//               the matrix combine A = cK*K + cM*M (+ an FP32 copy).
//  * spmv:      8 lanes per block row over the canonical upper-BSR values; the
//               row's own upper blocks and its transposed lower blocks
//               (through the transposed index) are one work list.
//  * PCG:       chunked vector kernels with per-instance FP64 reductions; the
//               last workgroup of every reduction kernel finalises the scalars
//               (alpha, beta, convergence) on the device.
//  * IC(0):     level-scheduled block factorisation and sweeps, one thread per row.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "od_elem.h"
#include "od_internal.h"

namespace od {

using namespace odk;

// =====================================================================================
// element kernel
// =====================================================================================
// The corotational stiffness block of node pair (a, b) at rest, K0_ab = V B_a^T D B_b for the
// isotropic D (D0 = lambda + 2 mu, D1 = lambda, D2 = mu) with factor = 1 / (36 V) and the
// gradients Ba, Bb scaled by 6 V (corotationalHyperelastic.cpp:20-41, tetelement.cpp:546-620).
__device__ __forceinline__ void corot_k0(const v3 Ba, const v3 Bb, double D0, double D1, double D2, double factor,
                                         double (&k)[9]) {
    k[0] = (D0 * Ba.x * Bb.x + D2 * (Ba.y * Bb.y + Ba.z * Bb.z)) * factor;
    k[1] = (D1 * Ba.x * Bb.y + D2 * Ba.y * Bb.x) * factor;
    k[2] = (D1 * Ba.x * Bb.z + D2 * Ba.z * Bb.x) * factor;
    k[3] = (D1 * Ba.y * Bb.x + D2 * Ba.x * Bb.y) * factor;
    k[4] = (D0 * Ba.y * Bb.y + D2 * (Ba.x * Bb.x + Ba.z * Bb.z)) * factor;
    k[5] = (D1 * Ba.y * Bb.z + D2 * Ba.z * Bb.y) * factor;
    k[6] = (D1 * Ba.z * Bb.x + D2 * Ba.x * Bb.z) * factor;
    k[7] = (D1 * Ba.z * Bb.y + D2 * Ba.y * Bb.z) * factor;
    k[8] = (D0 * Ba.z * Bb.z + D2 * (Ba.x * Bb.x + Ba.y * Bb.y)) * factor;
}

// KK = R k R^T
__device__ __forceinline__ void rotate_block(const double *R, const double *k, double *KK) {
    double QK[9];
    mul3(R, k, QK);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            KK[i * 3 + j] = QK[i * 3 + 0] * R[j * 3 + 0] + QK[i * 3 + 1] * R[j * 3 + 1] + QK[i * 3 + 2] * R[j * 3 + 2];
}

__device__ __forceinline__ void store_block(double *dst, const double *m, int transpose) {
    if (transpose) {
        dst[0] = m[0]; dst[1] = m[3]; dst[2] = m[6];
        dst[3] = m[1]; dst[4] = m[4]; dst[5] = m[7];
        dst[6] = m[2]; dst[7] = m[5]; dst[8] = m[8];
    } else {
#pragma unroll
        for (int i = 0; i < 9; i++) dst[i] = m[i];
    }
}

// Corotational plastic stress of tet e (corotationalPlastic.cpp:124-183): elastic strain
// sym(S) - I - plastic strain with S = R^T F, stress R (lambda tr(e) I + 2 mu e); then the
// plastic strain creeps towards the elastic strain when its norm exceeds the yield criterion
// and is capped at the maximum.  Both norms sum e[i]^2 over i, j < 3 with only i indexing the
// flattened strain, 3 (e00^2 + e01^2 + e02^2), as the reference's loops (:159, :172).
__device__ __forceinline__ void plastic_stress(const DevMesh &m, int e, const double *F, const double *R,
                                               double lambda, double mu, double *P) {
    double *ep = m.plStrain + 9 * (size_t)e;
    double S[9], st[9], ee[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i * 3 + j] = R[0 * 3 + i] * F[0 * 3 + j] + R[1 * 3 + i] * F[1 * 3 + j] + R[2 * 3 + i] * F[2 * 3 + j];
    double pl[9];
#pragma unroll
    for (int q = 0; q < 9; q++) pl[q] = ep[q];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) ee[i * 3 + j] = 0.5 * (S[i * 3 + j] + S[j * 3 + i]) - pl[i * 3 + j];
    ee[0] -= 1.0; ee[4] -= 1.0; ee[8] -= 1.0;
    const double mu2 = 2.0 * mu, dg = lambda * (ee[0] + ee[4] + ee[8]);
#pragma unroll
    for (int q = 0; q < 9; q++) st[q] = mu2 * ee[q];
    st[0] += dg; st[4] += dg; st[8] += dg;
    mul3(R, st, P);
    const double en = sqrt(3.0 * (ee[0] * ee[0] + ee[1] * ee[1] + ee[2] * ee[2]));
    if (en > m.plYield[e]) {
        const double ratio = m.plCreep[e] * (en - m.plYield[e]) / en;
#pragma unroll
        for (int q = 0; q < 9; q++) pl[q] += ratio * ee[q];
        const double pn = sqrt(3.0 * (pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]));
        if (pn > m.plMax[e]) {
            const double f = m.plMax[e] / pn;
#pragma unroll
            for (int q = 0; q < 9; q++) pl[q] *= f;
        }
#pragma unroll
        for (int q = 0; q < 9; q++) ep[q] = pl[q];
    }
}

// One kernel per material kind (KIND 0 corotational, 1 neo-Hookean, 2 corotational plastic)
// and one for the mass blocks (KIND 3): each compiles to its own register budget, and the
// budget allows two waves per SIMD.  A launch skips the tets of other kinds.
template <int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void k_element(
    DevMesh m, const double *__restrict__ d, const double *__restrict__ v, const double *__restrict__ a, ElemParams P,
    double *__restrict__ blkC, double *__restrict__ nodeC, const uint8_t *__restrict__ instActive) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.nTets) return;
    if (instActive && !instActive[m.tetInst[e]]) return;
    if (KIND != 3 && m.kind[e] != KIND) return;
    int tn[4], row[4];
    v3 X[4], x[4], y[4], z[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        tn[k] = m.tets[4 * e + k];
        row[k] = m.nodeRow[tn[k]];
        X[k] = mk(m.X[3 * tn[k]], m.X[3 * tn[k] + 1], m.X[3 * tn[k] + 2]);
        v3 u = mk(0, 0, 0), vv = mk(0, 0, 0), aa = mk(0, 0, 0);
        if (m.nodeDof) {   // partially constrained nodes: per-axis DOFs
            double c[3][3] = {};
#pragma unroll
            for (int ax = 0; ax < 3; ax++) {
                const int o = m.nodeDof[3 * tn[k] + ax];
                if (o < 0) continue;
                if (d) c[0][ax] = d[o];
                if (v) c[1][ax] = v[o];
                if (a) c[2][ax] = a[o];
            }
            u = mk(c[0][0], c[0][1], c[0][2]);
            vv = mk(c[1][0], c[1][1], c[1][2]);
            aa = mk(c[2][0], c[2][1], c[2][2]);
        } else if (row[k] >= 0) {
            int o = 3 * row[k];
            if (d) u = mk(d[o], d[o + 1], d[o + 2]);
            if (v) vv = mk(v[o], v[o + 1], v[o + 2]);
            if (a) aa = mk(a[o], a[o + 1], a[o + 2]);
        }
        x[k] = X[k] + u;
        y[k] = vv * P.cyv + aa * P.cya;
        z[k] = vv * P.czv + aa * P.cza;
    }
    const double V0 = tet_volume(X[0], X[1], X[2], X[3]);
    const double rho = m.rho ? m.rho[e] : 0.0;
    v3 g[4];
#pragma unroll
    for (int k = 0; k < 4; k++) g[k] = mk(0, 0, 0);

    if (KIND == 3) {  // consistent mass blocks rho*V/10 (diagonal pair) and rho*V/20 (tetelement.cpp:51-66)
        int p = 0;
        for (int aI = 0; aI < 4; aI++)
            for (int bI = aI; bI < 4; bI++, p++) {
                int slot = m.pairSlot[10 * e + p];
                if (slot < 0) continue;
                double mv = rho * (aI == bI ? V0 * 0.1 : V0 * 0.05);
                double *dst = blkC + (size_t)slot * 9;
                dst[0] = mv; dst[1] = 0; dst[2] = 0; dst[3] = 0; dst[4] = mv; dst[5] = 0; dst[6] = 0; dst[7] = 0; dst[8] = mv;
            }
        return;
    }

    v3 dN[4];
    tet_dn(X[0], X[1], X[2], X[3], dN);
    const double factor = 1.0 / (36.0 * V0);
    // rest shape matrix inverse and deformation gradient F = sum_i (x_i - x_0) (x) DmInv row (i-1)
    double Dm[9] = {X[1].x - X[0].x, X[2].x - X[0].x, X[3].x - X[0].x,
                    X[1].y - X[0].y, X[2].y - X[0].y, X[3].y - X[0].y,
                    X[1].z - X[0].z, X[2].z - X[0].z, X[3].z - X[0].z};
    double Di[9];
    inv3(Dm, Di);
    double F[9];
#pragma unroll
    for (int i = 0; i < 9; i++) F[i] = 0.0;
#pragma unroll
    for (int i = 1; i < 4; i++) {
        v3 ax = x[i] - x[0];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            F[0 + k] += ax.x * Di[(i - 1) * 3 + k];
            F[3 + k] += ax.y * Di[(i - 1) * 3 + k];
            F[6 + k] += ax.z * Di[(i - 1) * 3 + k];
        }
    }
    const double En = m.E[e], nu = m.nu[e];
    const double lambda = nu * En / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = En / (2.0 * (1.0 + nu));

    if (KIND == 0 || KIND == 2) {
        // ---- corotational: R from polar (or QR when compressed), blocks R K0_ab R^T
        double R[9];
        const double Vc = tet_volume(x[0], x[1], x[2], x[3]);
        if (Vc >= m.thrCorot * V0) polar3(F, R);
        else qr3(F, R);
        const double D0 = lambda + 2.0 * mu, D1 = lambda, D2 = mu;
        double Pst[9];
        if (KIND == 2) plastic_stress(m, e, F, R, lambda, mu, Pst);
        // corotational force plus the K_e y term of pair (a, b) (tetelement.cpp:622-647):
        // R k R^T (x_b + y_b) - R k X_b = R k w_b with w_b = R^T (x_b + y_b) - X_b, and for the
        // transposed pair R k^T w_a; the k w sums are rotated once per node at the end
        v3 w[4], h[4];
        if (KIND == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const v3 xy = x[q] + y[q];
                w[q] = mk(R[0] * xy.x + R[3] * xy.y + R[6] * xy.z, R[1] * xy.x + R[4] * xy.y + R[7] * xy.z,
                          R[2] * xy.x + R[5] * xy.y + R[8] * xy.z) - X[q];
                h[q] = mk(0, 0, 0);
            }
        }
        // compact form (stiffness pass): the rotated gradients g_q = R dN_q and the factors
        // lambda / (36 V0), mu / (36 V0); the block gather forms the blocks from them
        const bool compact = KIND == 0 && m.rotC != nullptr;
        if (compact) {
            double *o = m.rotC + 16 * (size_t)e;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const v3 gq = mulv(R, dN[q]);
                o[3 * q] = gq.x; o[3 * q + 1] = gq.y; o[3 * q + 2] = gq.z;
            }
            o[12] = lambda * factor;
            o[13] = mu * factor;
        }
        // both pair loops unrolled: every per-node array index is a constant (registers)
#pragma unroll
        for (int aI = 0; aI < 4; aI++) {
#pragma unroll
            for (int bI = aI; bI < 4; bI++) {
                const int p = 4 * aI - aI * (aI - 1) / 2 + (bI - aI);   // pair index (aI, bI), aI <= bI
                const v3 Ba = dN[aI], Bb = dN[bI];
                double k[9];
                corot_k0(Ba, Bb, D0, D1, D2, factor, k);
                double KK[9];
                if (!compact) {
                    rotate_block(R, k, KK);
                    int slot = m.pairSlot[10 * e + p];
                    if (slot >= 0) store_block(blkC + (size_t)slot * 9, KK, m.pairT[10 * e + p]);
                }
                if (!P.nodeTerms) continue;
                if (KIND == 2) {   // the K_e y term; the forces come from the elastic stress below
                    g[aI] = g[aI] + mulv(KK, y[bI]);
                    if (bI != aI) {
                        v3 ya = y[aI];
                        g[bI] = g[bI] + mk(KK[0] * ya.x + KK[3] * ya.y + KK[6] * ya.z,
                                           KK[1] * ya.x + KK[4] * ya.y + KK[7] * ya.z,
                                           KK[2] * ya.x + KK[5] * ya.y + KK[8] * ya.z);
                    }
                    continue;
                }
                h[aI] = h[aI] + mulv(k, w[bI]);
                if (bI != aI) {
                    const v3 wa = w[aI];
                    h[bI] = h[bI] + mk(k[0] * wa.x + k[3] * wa.y + k[6] * wa.z, k[1] * wa.x + k[4] * wa.y + k[7] * wa.z,
                                       k[2] * wa.x + k[5] * wa.y + k[8] * wa.z);
                }
            }
        }
        if (KIND == 0 && P.nodeTerms) {
#pragma unroll
            for (int q = 0; q < 4; q++) g[q] = mulv(R, h[q]);
        }
        if (KIND == 2 && P.nodeTerms) {
            // nodal forces from the elastic stress, f_i = P dN_i / 6, f_3 = -(f_0 + f_1 + f_2)
            // (corotationalPlastic.cpp:107-112, tetelement.cpp:715-729)
            const double f6 = 1.0 / 6.0;
            v3 f[3];
#pragma unroll
            for (int i = 0; i < 3; i++) f[i] = mulv(Pst, dN[i]) * f6;
            g[0] = g[0] + f[0];
            g[1] = g[1] + f[1];
            g[2] = g[2] + f[2];
            g[3] = g[3] - (f[0] + f[1] + f[2]);
        }
    } else {
        // ---- invertible neo-Hookean: P and the PSD-projected dP/dF in the diagonal frame
        double U[9], VT[9], Ps[9], tm[6], al[3], be[3];
        neo_stress(F, lambda, mu, m.thrNeo, U, VT, Ps, tm, al, be);
        v3 dr[4];
#pragma unroll
        for (int k = 0; k < 4; k++) dr[k] = mulv(VT, dN[k]);
        int p = 0;
        for (int aI = 0; aI < 4; aI++) {
            for (int bI = aI; bI < 4; bI++, p++) {
                const v3 na = dr[aI], nb = dr[bI];
                double S[9];
                S[0] = tm[0] * na.x * nb.x + al[0] * na.y * nb.y + al[1] * na.z * nb.z;
                S[4] = tm[3] * na.y * nb.y + al[0] * na.x * nb.x + al[2] * na.z * nb.z;
                S[8] = tm[5] * na.z * nb.z + al[1] * na.x * nb.x + al[2] * na.y * nb.y;
                S[1] = tm[1] * na.x * nb.y + be[0] * na.y * nb.x;
                S[3] = tm[1] * na.y * nb.x + be[0] * na.x * nb.y;
                S[2] = tm[2] * na.x * nb.z + be[1] * na.z * nb.x;
                S[6] = tm[2] * na.z * nb.x + be[1] * na.x * nb.z;
                S[5] = tm[4] * na.y * nb.z + be[2] * na.z * nb.y;
                S[7] = tm[4] * na.z * nb.y + be[2] * na.y * nb.z;
                double US[9], KK[9];
                mul3(U, S, US);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        KK[i * 3 + j] = (US[i * 3 + 0] * U[j * 3 + 0] + US[i * 3 + 1] * U[j * 3 + 1] + US[i * 3 + 2] * U[j * 3 + 2]) * factor;
                int slot = m.pairSlot[10 * e + p];
                if (slot >= 0) store_block(blkC + (size_t)slot * 9, KK, m.pairT[10 * e + p]);
                if (!P.nodeTerms) continue;
                g[aI] = g[aI] + mulv(KK, y[bI]);
                if (bI != aI) {
                    v3 ya = y[aI];
                    g[bI] = g[bI] + mk(KK[0] * ya.x + KK[3] * ya.y + KK[6] * ya.z, KK[1] * ya.x + KK[4] * ya.y + KK[7] * ya.z,
                                       KK[2] * ya.x + KK[5] * ya.y + KK[8] * ya.z);
                }
            }
        }
        if (P.nodeTerms) {
            // nodal forces f_i = P dN_i / 6, f_3 = -(f_0 + f_1 + f_2)  (tetelement.cpp:344-355)
            const double f6 = 1.0 / 6.0;
            v3 f[3];
#pragma unroll
            for (int i = 0; i < 3; i++) f[i] = mulv(Ps, dN[i]) * f6;
            g[0] = g[0] + f[0];
            g[1] = g[1] + f[1];
            g[2] = g[2] + f[2];
            g[3] = g[3] - (f[0] + f[1] + f[2]);
        }
    }
    if (!P.nodeTerms) return;
    if (P.cm != 0.0) {
        const double md = rho * V0 * 0.1, mo = rho * V0 * 0.05;
        v3 zs = z[0] + z[1] + z[2] + z[3];
#pragma unroll
        for (int k = 0; k < 4; k++) g[k] = g[k] + (z[k] * (md - mo) + zs * mo) * P.cm;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int slot = m.nodeSlot[4 * e + k];
        if (slot < 0) continue;
        double *dst = nodeC + (size_t)slot * 3;
        dst[0] = g[k].x; dst[1] = g[k].y; dst[2] = g[k].z;
    }
}

void launch_element(const DevMesh &m, hipStream_t s, const double *d, const double *v, const double *a,
                    const ElemParams &p, double *blkC, double *nodeC, const uint8_t *instActive) {
    int grid = (m.nTets + 127) / 128;
    if (p.mode == 1) {
        hipLaunchKernelGGL(k_element<3>, dim3(grid), dim3(128), 0, s, m, d, v, a, p, blkC, nodeC, instActive);
    } else {
        if (m.hasKind[0])
            hipLaunchKernelGGL(k_element<0>, dim3(grid), dim3(128), 0, s, m, d, v, a, p, blkC, nodeC, instActive);
        if (m.hasKind[1])
            hipLaunchKernelGGL(k_element<1>, dim3(grid), dim3(128), 0, s, m, d, v, a, p, blkC, nodeC, instActive);
        if (m.hasKind[2])
            hipLaunchKernelGGL(k_element<2>, dim3(grid), dim3(128), 0, s, m, d, v, a, p, blkC, nodeC, instActive);
    }
    OD_HIP(hipGetLastError());
}

// =====================================================================================
// gathers
// =====================================================================================
__constant__ int c_diagSrc[6] = {0, 1, 2, 4, 5, 8};

// One thread per stored value, its contributions in slot order.  SKIPC: a stiffness pass on a
// mesh with corotational tets in compact form and tets of other kinds, whose blocks with a
// corotational contribution the compact gather forms; this pass takes the others.
template <bool SKIPC>
__global__ void k_gather_blocks(DevMesh m, const double *__restrict__ C, double cK, const double *__restrict__ M,
                                double cM, double *__restrict__ out, int accumulate, float *__restrict__ out32,
                                const uint8_t *__restrict__ instActive) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= m.nBlocks * 9) return;
    const int b = (int)(gid / 9), c = (int)(gid % 9);
    if (SKIPC && m.blkCompact[b]) return;
    const int r = m.blkRow[b], col = m.colIdx[b], k0 = m.blkCStart[b], k1 = m.blkCStart[b + 1];
    const int64_t bo = m.blkOff[b];
    if (instActive && !instActive[m.rowInst[r]]) return;
    const bool diag = col == r;
    if (diag && c >= 6) return;
    const int src = diag ? c_diagSrc[c] : c;
    const int64_t vi = bo + c;
    // contributions in batches of 8 loads issued together; the sum keeps its order
    double s = 0.0;
    for (int k = k0; k < k1; k += 8) {
        double t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = k + j < k1 ? C[(size_t)(k + j) * 9 + src] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (k + j < k1) s += t[j];
    }
    double o = cK * s;
    if (M) o += cM * M[vi];
    if (accumulate) o += out[vi];
    out[vi] = o;
    if (out32) out32[vi] = (float)o;
}

// Partially constrained nodes: one thread per stored value, summing its element entries in
// tet order (the reference's scatter order); a padding diagonal sums 1.0, other padding
// values nothing.
__global__ void k_gather_values(DevMesh m, const double *__restrict__ C, double cK, const double *__restrict__ M,
                                double cM, double *__restrict__ out, int accumulate, float *__restrict__ out32,
                                const uint8_t *__restrict__ instActive) {
    const int64_t vi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (vi >= m.nValues) return;
    if (instActive && !instActive[0]) return;   // one instance (partially constrained meshes)
    const int k0 = m.vPtr[vi], k1 = m.vPtr[vi + 1];
    double s = 0.0;
    for (int k = k0; k < k1; k++) {
        const int src = m.vSrc[k];
        s += src < 0 ? 1.0 : C[src];
    }
    double o = cK * s;
    if (M) o += cM * M[vi];
    if (accumulate) o += out[vi];
    out[vi] = o;
    if (out32) out32[vi] = (float)o;
}

// Node pair (a, b), a <= b, of pair index p (0-9).
__constant__ int c_pairA[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
__constant__ int c_pairB[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};

// A = cK K + cM M with the compact corotational form.  A corotational tet's block is
// R K0_ab R^T; with K0_ab = f (lambda a b^T + mu b a^T + mu (a . b) I) (the isotropic B^T D B
// of corot_k0, a = dN_a, b = dN_b, f = 1 / (36 V0)) and R R^T = I that is
// f (lambda g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I) with the rotated gradients g = R dN
// the element pass stored; its transpose swaps a and b.  Any other tet's block is read from C.
// A block's contributions are cut, in slot order, into segments of at most kGatherSeg; each
// segment is summed in slot order by one thread and the segment sums are added in segment
// order, so the rounding is fixed by the mesh, not by the launch.
template <bool ALLCOROT>
__global__ void k_gather_blocks_rot(DevMesh m, const double *__restrict__ C, double cK, const double *__restrict__ M,
                                    double cM, double *__restrict__ out, int accumulate, float *__restrict__ out32,
                                    const uint8_t *__restrict__ instActive) {
    // blocks in the SpMV's run order (chunks sorted in space, one run per XCD): the j-th
    // workgroup on XCD x takes the next 256 segments of run x, so an XCD's L2 keeps the tets of
    // its slab while all their blocks are formed.  A thread sums one segment (at most
    // kGatherSeg contributions); the block's first segment adds the others' sums in order.
    const int j = (int)(blockIdx.x >> 3);
    if (j >= m.gatherPerRun) return;   // the whole workgroup
    const int4 sg = m.gatherSeg[((int64_t)(blockIdx.x & 7) * m.gatherPerRun + j) * kBlock + threadIdx.x];
    const int b = sg.x, k0 = sg.y, k1 = sg.z;
    bool act = b >= 0 && (ALLCOROT || m.blkCompact[b]);   // else padding, or the stored-block pass's
    int r = 0, col = 0;
    if (act) {
        r = m.blkRow[b];
        col = m.colIdx[b];
        if (instActive && !instActive[m.rowInst[r]]) act = false;
    }
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // contributions in batches of four whose loads are issued together (past the end: slot k0,
    // dropped); the sums keep the slot order.  A corotational contribution loads g_a, g_b and
    // the two factors, any other its stored block (the source's stored flag, set from the
    // tets' kinds, picks which)
    constexpr int NB = kGatherSeg < 4 ? kGatherSeg : 4;
    for (int k = k0; act && k < k1; k += NB) {
        int src[NB];
        double gv[NB][9];
#pragma unroll
        for (int j = 0; j < NB; j++) src[j] = m.slotSrc[k + j < k1 ? k + j : k0];
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (ALLCOROT || !(src[j] & kSrcStored)) {
                const int e = src[j] >> kSrcTetShift, p = (src[j] >> 1) & 15, tr = src[j] & 1;
                const int a = tr ? c_pairB[p] : c_pairA[p], bb = tr ? c_pairA[p] : c_pairB[p];
                const double *g = m.rotC + 16 * (size_t)e;
                gv[j][0] = g[3 * a]; gv[j][1] = g[3 * a + 1]; gv[j][2] = g[3 * a + 2];
                gv[j][3] = g[3 * bb]; gv[j][4] = g[3 * bb + 1]; gv[j][5] = g[3 * bb + 2];
                gv[j][6] = g[12]; gv[j][7] = g[13];
            } else {
                const double *c = C + (size_t)(k + j < k1 ? k + j : k0) * 9;
#pragma unroll
                for (int i = 0; i < 9; i++) gv[j][i] = c[i];
            }
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (k + j >= k1) break;
            double kk[9];
            if (ALLCOROT || !(src[j] & kSrcStored)) {
                const double *ga = gv[j], *gb = gv[j] + 3;
                const double lf = gv[j][6], mf = gv[j][7];
                const double dm = mf * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int q = 0; q < 3; q++) kk[3 * i + q] = lf * ga[i] * gb[q] + mf * gb[i] * ga[q] + (i == q ? dm : 0.0);
            } else {
#pragma unroll
                for (int i = 0; i < 9; i++) kk[i] = gv[j][i];
            }
#pragma unroll
            for (int i = 0; i < 9; i++) s[i] += kk[i];
        }
    }
    // the block's segments combine in order through LDS
    __shared__ double part[kBlock][9];
#pragma unroll
    for (int i = 0; i < 9; i++) part[threadIdx.x][i] = s[i];
    __syncthreads();
    if (!act || sg.w < 0) return;
    for (int q = 1; q <= sg.w; q++)
#pragma unroll
        for (int i = 0; i < 9; i++) s[i] += part[threadIdx.x + q][i];
    const int64_t bo = m.blkOff[b];
    const bool diag = col == r;
    const int nv = diag ? 6 : 9;
    for (int c = 0; c < nv; c++) {
        const int64_t vi = bo + c;
        double o = cK * s[diag ? c_diagSrc[c] : c];
        if (M) o += cM * M[vi];
        if (accumulate) o += out[vi];
        out[vi] = o;
        if (out32) out32[vi] = (float)o;
    }
}

// The stored flag of every contribution source from its tet's material kind (set with the
// materials): the compact gather reads the stored block of any tet that is not corotational.
__global__ void k_mark_stored(int *slotSrc, int64_t n, const int *__restrict__ kind) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int sv = slotSrc[k];
    slotSrc[k] = (sv & ~kSrcStored) | (kind[sv >> kSrcTetShift] != 0 ? kSrcStored : 0);
}

// Per block, whether any of its contributions is corotational (formed by the compact gather).
__global__ void k_mark_compact_blocks(DevMesh m) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nBlocks) return;
    uint8_t f = 0;
    for (int k = m.blkCStart[b]; k < m.blkCStart[b + 1]; k++)
        if (!(m.slotSrc[k] & kSrcStored)) f = 1;
    m.blkCompact[b] = f;
}

void launch_mark_stored(const DevMesh &m, hipStream_t s, int64_t nContrib) {
    if (!m.slotSrc || nContrib <= 0) return;
    hipLaunchKernelGGL(k_mark_stored, dim3((unsigned)((nContrib + 255) / 256)), dim3(256), 0, s, m.slotSrc, nContrib,
                       m.kind);
    OD_HIP(hipGetLastError());
    if (m.blkCompact) {
        hipLaunchKernelGGL(k_mark_compact_blocks, dim3((unsigned)((m.nBlocks + 255) / 256)), dim3(256), 0, s, m);
        OD_HIP(hipGetLastError());
    }
}

void launch_gather_blocks(const DevMesh &m, hipStream_t s, const double *blkC, double cK, const double *M, double cM,
                          double *out, int accumulate, float *out32, const uint8_t *instActive, bool stiff) {
    if (stiff && m.rotC && m.hasKind[0] && !m.vPtr) {
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(8 * (unsigned)m.gatherPerRun), dim3(kBlock), 0, s, m, blkC, cK, M, cM, out,
                               accumulate, out32, instActive);
        };
        if (!m.hasKind[1] && !m.hasKind[2]) {
            go(k_gather_blocks_rot<true>);
        } else {
            // other kinds too: their blocks value by value (coalesced reads of the stored
            // blocks), then the blocks with a corotational contribution in compact form
            const int64_t n = m.nBlocks * 9;
            hipLaunchKernelGGL(k_gather_blocks<true>, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, m,
                               blkC, cK, M, cM, out, accumulate, out32, instActive);
            go(k_gather_blocks_rot<false>);
        }
        OD_HIP(hipGetLastError());
        return;
    }
    if (m.vPtr) {
        hipLaunchKernelGGL(k_gather_values, dim3((unsigned)((m.nValues + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, m,
                           blkC, cK, M, cM, out, accumulate, out32, instActive);
        OD_HIP(hipGetLastError());
        return;
    }
    int64_t n = m.nBlocks * 9;
    int grid = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_gather_blocks<false>, dim3(grid), dim3(kBlock), 0, s, m, blkC, cK, M, cM, out, accumulate,
                       out32, instActive);
    OD_HIP(hipGetLastError());
}

// mode 0: out = cf*(fext + imp - sum);  mode 1: out += sum (internal force accumulate)
__global__ void k_gather_rhs(DevMesh m, const double *__restrict__ C, const double *__restrict__ fext,
                             const double *__restrict__ imp, double cf, double *__restrict__ out, int mode,
                             const uint8_t *__restrict__ instActive) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * m.nRows) return;
    const int r = i / 3, ax = i % 3;
    if (instActive && !instActive[m.rowInst[r]]) return;
    const int k0 = m.rowCStart[r], k1 = m.rowCStart[r + 1];
    double s = 0.0;
    for (int k = k0; k < k1; k += 8) {
        double t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = k + j < k1 ? C[(size_t)(k + j) * 3 + ax] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (k + j < k1) s += t[j];
    }
    if (mode == 1) { out[i] += s; return; }
    double f = fext ? fext[i] : 0.0;
    if (imp) f += imp[i];
    out[i] = cf * (f - s);
}

// Partially constrained nodes: per DOF, its node-force entries in tet order (padding DOFs have
// none and keep fext = imp = 0).
__global__ void k_gather_rhs_dofs(DevMesh m, const double *__restrict__ C, const double *__restrict__ fext,
                                  const double *__restrict__ imp, double cf, double *__restrict__ out, int mode,
                                  const uint8_t *__restrict__ instActive) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * m.nRows) return;
    if (instActive && !instActive[0]) return;   // one instance (partially constrained meshes)
    double s = 0.0;
    for (int k = m.dPtr[i]; k < m.dPtr[i + 1]; k++) s += C[m.dSrc[k]];
    if (mode == 1) { out[i] += s; return; }
    double f = fext ? fext[i] : 0.0;
    if (imp) f += imp[i];
    out[i] = cf * (f - s);
}

void launch_gather_rhs(const DevMesh &m, hipStream_t s, const double *nodeC, const double *fext, const double *imp,
                       double cf, double *out, int mode, const uint8_t *instActive) {
    int n = 3 * m.nRows;
    if (m.dPtr) {
        hipLaunchKernelGGL(k_gather_rhs_dofs, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m, nodeC, fext, imp,
                           cf, out, mode, instActive);
        OD_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_gather_rhs, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m, nodeC, fext, imp, cf, out,
                       mode, instActive);
    OD_HIP(hipGetLastError());
}

// body force V/4 * b per node (tetelement.cpp:68-77), written per node slot
__global__ void k_body(DevMesh m, const double *__restrict__ body, double *__restrict__ nodeC) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.nTets) return;
    v3 X[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int n = m.tets[4 * e + k];
        X[k] = mk(m.X[3 * n], m.X[3 * n + 1], m.X[3 * n + 2]);
    }
    double f = tet_volume(X[0], X[1], X[2], X[3]) / 4.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int slot = m.nodeSlot[4 * e + k];
        if (slot < 0) continue;
        nodeC[(size_t)slot * 3 + 0] = f * body[3 * e + 0];
        nodeC[(size_t)slot * 3 + 1] = f * body[3 * e + 1];
        nodeC[(size_t)slot * 3 + 2] = f * body[3 * e + 2];
    }
}

void launch_body_force(const DevMesh &m, hipStream_t s, const double *bodyPerTet, double *nodeC, double *f) {
    hipLaunchKernelGGL(k_body, dim3((m.nTets + 127) / 128), dim3(128), 0, s, m, bodyPerTet, nodeC);
    OD_HIP(hipGetLastError());
    OD_HIP(hipMemsetAsync(f, 0, sizeof(double) * 3 * (size_t)m.nRows, s));
    launch_gather_rhs(m, s, nodeC, nullptr, nullptr, 1.0, f, 1, nullptr);
}

// Expands the block-level map into the reference's 78-entry per-tet map.
__global__ void k_expand_map(DevMesh m, int *__restrict__ map78) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.nTets) return;
    int *out = map78 + (size_t)78 * e;
    // packed order: sub-column c = 3a+i, sub-row r >= c (element.h:109-113)
    for (int c = 0; c < 12; c++) {
        for (int r = c; r < 12; r++) {
            int li = c * 12 - ((c - 1) * c) / 2 + (r - c);
            int aN = c / 3, i = c % 3, bN = r / 3, j = r % 3;   // aN <= bN
            int p = aN * 4 - (aN * (aN - 1)) / 2 + (bN - aN);  // pair index of (aN, bN), aN <= bN
            int b = m.pairBlk[10 * e + p];
            if (b < 0) { out[li] = -1; continue; }
            int base = m.blkOff[b];
            if (aN == bN) {
                // packed upper row-major of the diagonal block, entry (i, j), i <= j
                out[li] = base + (i == 0 ? j : (i == 1 ? 2 + j : 5));
            } else {
                // element entry (row a-DOF i, col b-DOF j); stored block is (lowRow, highRow) row-major
                int t = m.pairT[10 * e + p];
                out[li] = base + (t ? (j * 3 + i) : (i * 3 + j));
            }
        }
    }
}

void launch_expand_map(const DevMesh &m, hipStream_t s, int *map78) {
    hipLaunchKernelGGL(k_expand_map, dim3((m.nTets + 127) / 128), dim3(128), 0, s, m, map78);
    OD_HIP(hipGetLastError());
}

// =====================================================================================
// reductions: per-chunk partials, then one finalisation workgroup per instance
// =====================================================================================
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums NQ per-thread values over the workgroup in a fixed order; thread 0 gets the result.
template <int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ]) {
    static_assert(NQ >= 1 && NQ <= 4, "block_sum holds up to 4 values per wave");
    __shared__ double lds[16 * 4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double s = wave_sum(v[q]);
        if (lane == 0) lds[w * NQ + q] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double s = 0;
            for (int k = 0; k < nw; k++) s += lds[k * NQ + q];
            v[q] = s;
        }
    }
}

// partial slots per chunk: [0] r.z or p.q, [1] r.r, [2] b.b
enum Phase { PH_INIT = 0, PH_ALPHA = 1, PH_ITER = 2, PH_OUTER_FIRST = 3, PH_OUTER = 4 };

struct FinalArgs {
    double tol;
    int maxIt;
    const uint8_t *preDone;
    const InstScalars *outer;
};

// An inner (FP32) solve of the mixed form stops at innerTol relative to its own right-hand
// side, or as soon as its residual is below half the outer FP64 tolerance, whichever is
// looser: the correction then meets the outer test without a second sweep.
constexpr double kInnerOuterScale2 = 0.25;

// Sums the chunk partials of slots [Q0, Q0 + NQ) over chunks [c0, c1) in the order of 1024
// threads whatever the workgroup size (a power of two from 64 to 1024): virtual thread v sums
// chunks v, v + 1024, ... in order, each virtual wave of 64 is summed by wave_sum, and
// thread 0 adds the 16 wave sums in order.  So an instance finalized by the 256-thread
// workgroups of a batch gets the bits of the same instance alone (1024 threads), whatever
// its size.  Only the slots a phase uses are read.
template <int Q0, int NQ>
__device__ __forceinline__ void reduce_slots(const DevMesh &m, const double *__restrict__ partials, int c0, int c1,
                                             double *t) {
    constexpr int kV = 1024;
    __shared__ double lds[(kV / 64) * NQ];
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int j = 0; j < kV / (int)blockDim.x; j++) {
        const int vt = (int)threadIdx.x + j * (int)blockDim.x;
        double v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) v[q] = 0.0;
#pragma unroll 8
        for (int c = c0 + vt; c < c1; c += kV) {
#pragma unroll
            for (int q = 0; q < NQ; q++) v[q] += partials[(Q0 + q) * m.nChunks + c];
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const double sw = wave_sum(v[q]);
            if (lane == 0) lds[((int)(threadIdx.x >> 6) + j * nw) * NQ + q] = sw;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double sum = 0;
            for (int k = 0; k < kV / 64; k++) sum += lds[k * NQ + q];
            t[Q0 + q] = sum;
        }
    }
}

// Reduces an instance's chunk partials (fixed order) and updates its scalars; called by every
// thread of a workgroup (any size up to 1024).
__device__ void finalize_inst(const DevMesh &m, int phase, int inst, const double *__restrict__ partials,
                              InstScalars *__restrict__ sc, int *__restrict__ doneFlags, const FinalArgs &fa) {
    // one instance: all chunks (no range load); the scalars are loaded with the partials
    const bool single = m.nInst == 1;
    const int c0 = single ? 0 : m.instChunkStart[inst], c1 = single ? m.nChunks : m.instChunkStart[inst + 1];
    InstScalars S{};
    if (threadIdx.x == 0) S = sc[inst];
    double t[3] = {0, 0, 0};
    if (phase == PH_ALPHA) reduce_slots<0, 1>(m, partials, c0, c1, t);
    else if (phase == PH_ITER) reduce_slots<0, 2>(m, partials, c0, c1, t);
    else if (phase == PH_INIT) reduce_slots<0, 3>(m, partials, c0, c1, t);
    else if (phase == PH_OUTER_FIRST) reduce_slots<1, 2>(m, partials, c0, c1, t);
    else reduce_slots<1, 1>(m, partials, c0, c1, t);
    if (threadIdx.x != 0) return;
    switch (phase) {
    case PH_INIT:
        S.delta = t[0]; S.rr = t[1]; S.bb = t[2];
        S.beta = 0.0;                       // the first product takes p = z
        S.it = 0; S.maxIt = fa.maxIt;
        S.eps = fa.tol * fa.tol * S.bb;
        if (fa.outer) S.eps = fmax(S.eps, kInnerOuterScale2 * fa.outer[inst].eps);
        S.done = (fa.preDone && fa.preDone[inst]) || !(S.rr > S.eps) || fa.maxIt <= 0;
        break;
    case PH_ALPHA:
        if (!S.done) { S.pq = t[0]; S.alpha = S.delta / t[0]; }
        break;
    case PH_ITER:
        if (!S.done) {
            S.rz = t[0]; S.rr = t[1];
            S.beta = S.rz / S.delta;
            S.delta = S.rz;
            S.it += 1;
            S.done = !(S.rr > S.eps) || S.it >= S.maxIt;
        }
        break;
    case PH_OUTER_FIRST:
    case PH_OUTER:
        S.rr = t[1];
        if (phase == PH_OUTER_FIRST) S.bb = t[2];
        S.eps = fa.tol * fa.tol * S.bb;
        S.done = (fa.preDone && fa.preDone[inst]) || !(S.rr > S.eps);
        break;
    }
    sc[inst] = S;
    doneFlags[inst] = S.done;
}

// One workgroup per instance.
__global__ __launch_bounds__(1024) void k_finalize(DevMesh m, int phase, const double *__restrict__ partials,
                                                  InstScalars *__restrict__ sc, int *__restrict__ doneFlags,
                                                  FinalArgs fa) {
    finalize_inst(m, phase, blockIdx.x, partials, sc, doneFlags, fa);
}

void launch_finalize(const DevMesh &m, hipStream_t s, int phase, PcgWork &w, double tol, int maxIt,
                     const uint8_t *preDone) {
    // a wide workgroup for few large instances, a narrow one for many small instances
    const int threads = m.nInst == 1 ? 1024 : 256;
    hipLaunchKernelGGL(k_finalize, dim3(m.nInst), dim3(threads), 0, s, m, phase, w.partials, w.sc, w.doneFlags,
                       FinalArgs{tol, maxIt, preDone, phase == PH_INIT ? w.outer : nullptr});
    OD_HIP(hipGetLastError());
}

// =====================================================================================
// SpMV over the canonical upper-BSR layout: one workgroup = one 64-row chunk, 8 lanes per row
// =====================================================================================
// One stored 3x3 block and the three entries of x it multiplies.  Value arrays are padded
// by kValuePad entries so the 9-value load of a trailing diagonal block stays in bounds.
template <typename T>
struct BlockX {
    T e[9], x[3];
};

template <typename T>
__device__ __forceinline__ void load_block(const T *__restrict__ A, const T *__restrict__ xc, BlockX<T> &b) {
#pragma unroll
    for (int q = 0; q < 9; q++) b.e[q] = A[q];
    b.x[0] = xc[0]; b.x[1] = xc[1]; b.x[2] = xc[2];
}

// The PCG direction at one column, p = z + beta p_old (b.x holds p_old on entry): every
// workgroup that gathers a column computes it with the same fused multiply-add, so all
// copies of p_j agree to the bit.
template <typename T>
__device__ __forceinline__ void dir_update(const T *__restrict__ zc, T beta, BlockX<T> &b) {
    const T z0 = zc[0], z1 = zc[1], z2 = zc[2];
    b.x[0] = fma(beta, b.x[0], z0);
    b.x[1] = fma(beta, b.x[1], z1);
    b.x[2] = fma(beta, b.x[2], z2);
}

// One block-row term and its accumulation, rounded the same way in every kernel: the row
// dot product e0 x0 + e1 x1 + e2 x2 as explicit fused multiply-adds (x0 first), then one
// plain add into the running sum.  Contraction is off here so the compiler cannot fold the
// running sum into the products (a fold it makes or not depending on the loop's shape).
template <typename T>
__device__ __forceinline__ T row3(T e0, T e1, T e2, T x0, T x1, T x2) {
#pragma clang fp contract(off)
    return fma(e2, x2, fma(e1, x1, e0 * x0));
}
template <typename T>
__device__ __forceinline__ void acc_add(T &a, T t) {
#pragma clang fp contract(off)
    a = a + t;
}

// a += B x for the packed symmetric diagonal [e0 e1 e2; e1 e3 e4; e2 e4 e5], a row-major
// upper block, or the transpose of one.
template <typename T>
__device__ __forceinline__ void block_mv(const BlockX<T> &b, bool diag, bool trans, T &a0, T &a1, T &a2) {
    T e0 = b.e[0], e1 = b.e[1], e2 = b.e[2], e3 = b.e[3], e4 = b.e[4], e5 = b.e[5], e6 = b.e[6], e7 = b.e[7],
      e8 = b.e[8];
    if (diag) {
        e8 = e5; e7 = e4; e6 = e2; e5 = e4; e4 = e3; e3 = e1;
    } else if (trans) {
        T t;
        t = e1; e1 = e3; e3 = t;
        t = e2; e2 = e6; e6 = t;
        t = e5; e5 = e7; e7 = t;
    }
    acc_add(a0, row3(e0, e1, e2, b.x[0], b.x[1], b.x[2]));
    acc_add(a1, row3(e3, e4, e5, b.x[0], b.x[1], b.x[2]));
    acc_add(a2, row3(e6, e7, e8, b.x[0], b.x[1], b.x[2]));
}

template <int L, typename T>
__device__ __forceinline__ void sum_lanes(T &a0, T &a1, T &a2) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, L);
        a1 += __shfl_xor(a1, o, L);
        a2 += __shfl_xor(a2, o, L);
    }
}

// Lanes per row L (4, 8 or 16) is a launch choice only.  Each lane keeps the sums of 16 / L
// virtual lanes (v = lane + g L, which take blocks v, v + 16, ... of the row) and combines
// them in the order of a 16-lane shuffle tree (pairs at distance 8, then 4, 2, 1), so every
// L gives the bits of 16 lanes.
template <int L, typename T>
__device__ __forceinline__ void combine16(const T (&acc)[16 / L][3], T &a0, T &a1, T &a2) {
    if constexpr (L == 16) {
        a0 = acc[0][0]; a1 = acc[0][1]; a2 = acc[0][2];
    } else if constexpr (L == 8) {
        a0 = acc[0][0] + acc[1][0]; a1 = acc[0][1] + acc[1][1]; a2 = acc[0][2] + acc[1][2];
    } else {
        static_assert(L == 4, "4, 8 or 16 lanes per row");
        a0 = (acc[0][0] + acc[2][0]) + (acc[1][0] + acc[3][0]);
        a1 = (acc[0][1] + acc[2][1]) + (acc[1][1] + acc[3][1]);
        a2 = (acc[0][2] + acc[2][2]) + (acc[1][2] + acc[3][2]);
    }
    sum_lanes<L>(a0, a1, a2);
}

// One block row through the row pointers (rows with more than 16 blocks): item 0 is the
// diagonal block, items 1..nOff the row's upper blocks (contiguous values), the remaining
// items the transposed lower blocks reached through the transposed index.  LS lanes per row,
// each with the sums of 16 / LS virtual lanes (combine16: the bits of 16 lanes).
// With FUSE, x is p_old and each gathered column is the direction z + beta p_old; the lane
// of item 0 (lane 0) returns the row's own direction in px.
template <typename T, bool FUSE, int LS>
__device__ __forceinline__ void row_spmv(const DevMesh &m, const T *__restrict__ vals, const T *__restrict__ x, int r,
                                         int lane, T &o0, T &o1, T &o2, const T *__restrict__ z, T beta, T *px) {
    constexpr int G = kSpmvLanes / LS;
    const int v0 = m.valOff[r], b0 = m.rowPtr[r];
    const int nOff = m.rowPtr[r + 1] - b0 - 1;
    const int t0 = m.tPtr[r], nItems = 1 + nOff + (m.tPtr[r + 1] - t0);
    T acc[G][3];
#pragma unroll
    for (int g = 0; g < G; g++) acc[g][0] = acc[g][1] = acc[g][2] = 0;
    for (int it0 = lane; it0 < nItems; it0 += kSpmvLanes) {
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int it = it0 + g * LS;
            if (it >= nItems) continue;
            const bool up = it <= nOff;
            const int k = t0 + (up ? 0 : it - nOff - 1);
            const int col = up ? m.colIdx[b0 + it] : m.tRow[k];
            const T *A = up ? vals + (it == 0 ? v0 : v0 + 6 + 9 * (it - 1)) : vals + m.tOff[k];
            BlockX<T> b;
            load_block<T>(A, x + 3 * col, b);
            if (FUSE) {
                dir_update<T>(z + 3 * col, beta, b);
                if (it == 0) { px[0] = b.x[0]; px[1] = b.x[1]; px[2] = b.x[2]; }
            }
            block_mv<T>(b, it == 0, !up, acc[g][0], acc[g][1], acc[g][2]);
        }
    }
    combine16<LS>(acc, o0, o1, o2);
}

// Run position of the q-th (0-based) chunk of the run of the XCD this workgroup runs on, or
// -1 past the run's end: workgroup w runs on XCD (w mod 8) and is that XCD's (w div 8)-th.
__device__ __forceinline__ int run_pos(const DevMesh &m, int q) {
    const int x = (int)(blockIdx.x & 7), idx = x * m.perRun + q;
    return (q < m.perRun && idx < m.nChunks) ? idx : -1;
}

// y = A x (or y += A x) for R chunks per workgroup, LS lanes per block row (16, 8 or 4; 64 LS
// threads).  A row has 16 slots; lane l takes slots l, l + LS, ... and the row sum combines
// them in the order of a 16-lane tree (combine16), so every launch shape gives the same bits.
// Every load is unconditional (slots of padding, inactive or long-row chunks point at valid
// memory and their results are dropped), so the slot loads of all R chunks, then all their
// block and x loads, are in flight together.  The dot partial x.y of each chunk (sum over its
// 64 rows by one wave, a fixed order) goes to partials[chunk] when partials is set.
//
// With FUSE (the PCG product), x is the previous direction p_old: every gathered column is
// the new direction p = z + beta p_old of its instance (dir_update), the row's own direction
// (from its diagonal slot, lane 0) is written to pNew, and the dot partial is p.q.
//
// Dependent loads: the run record and the slots need only the workgroup index, and the block
// and x loads only the slots.  A mesh of one instance checks its converged flag first, from
// a fixed address (the whole grid exits together past convergence); with several instances
// a chunk of a converged instance loads nothing but its slots.
template <typename T, int R, bool FUSE, int LS>
__device__ __forceinline__ void spmv_chunks(const DevMesh &m, const T *__restrict__ vals, const T *__restrict__ x,
                                            T *__restrict__ y, int accumulate, double *__restrict__ partials,
                                            const InstScalars *__restrict__ sc, const T *__restrict__ z,
                                            T *__restrict__ pNew) {
    constexpr int G = kSpmvLanes / LS;
    const bool single = m.nInst == 1;
    if (single && sc && sc[0].done) return;
    const int j = (int)(blockIdx.x >> 3), lane = (int)threadIdx.x % LS, lr = (int)threadIdx.x / LS;
    int chunk[R], r[R], r1[R], off[R][G], cf[R][G];
    bool act[R], lng[R], use[R];
    T beta[R];
    int4 info[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int pos = run_pos(m, R * j + i), p = pos >= 0 ? pos : 0;
        act[i] = pos >= 0;
#pragma unroll
        for (int g = 0; g < G; g++) {
            const size_t slot = (size_t)p * kSpmvThreads + (size_t)lr * kSpmvLanes + lane + g * LS;
            off[i][g] = m.slotOff[slot];
            cf[i][g] = m.slotCol[slot];
        }
        info[i] = m.runInfo[p];
    }
#pragma unroll
    for (int i = 0; i < R; i++) {
        chunk[i] = act[i] ? info[i].x : -1;
        r[i] = info[i].y + lr;
        r1[i] = info[i].z;
        lng[i] = (info[i].w & kRunLong) != 0;
        const int inst = info[i].w & kRunInstMask;
        if (!single && sc && act[i]) act[i] = !sc[inst].done;
        use[i] = act[i] && !lng[i];          // the slots hold the whole rows
        beta[i] = FUSE ? (T)sc[single ? 0 : inst].beta : (T)0;
    }
    // slots not used (padding, converged instance, row-loop chunk) read x[0] and the first
    // value and count as padding (cf = -1) from here on
    BlockX<T> b[R][G];
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int g = 0; g < G; g++) {
            const bool ok = use[i] && cf[i][g] >= 0;
            if (!ok) cf[i][g] = -1;
            const int c3 = 3 * (ok ? (cf[i][g] & kSlotColMask) : 0);
            load_block<T>(vals + (ok ? off[i][g] : 0), x + c3, b[i][g]);
            if (FUSE) dir_update<T>(z + c3, beta[i], b[i][g]);
        }
    T acc[R][G][3];
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int g = 0; g < G; g++) {
            acc[i][g][0] = acc[i][g][1] = acc[i][g][2] = 0;
            block_mv<T>(b[i][g], (cf[i][g] & kSlotDiag) != 0, (cf[i][g] & kSlotTrans) != 0, acc[i][g][0],
                        acc[i][g][1], acc[i][g][2]);
            if (cf[i][g] < 0) acc[i][g][0] = acc[i][g][1] = acc[i][g][2] = 0;
        }
    __shared__ double rowDot[R][kChunkRows];
    // row output (lane 0): y (or y +=), the new direction with FUSE, and the row's dot term
    auto finish_row = [&](int i, T a0, T a1, T a2, const T *px) {
        const int rr = r[i];
        T d0 = px[0], d1 = px[1], d2 = px[2];
        if (!FUSE) { d0 = x[3 * rr]; d1 = x[3 * rr + 1]; d2 = x[3 * rr + 2]; }
        if (FUSE) { pNew[3 * rr] = d0; pNew[3 * rr + 1] = d1; pNew[3 * rr + 2] = d2; }
        if (accumulate) { a0 += y[3 * rr]; a1 += y[3 * rr + 1]; a2 += y[3 * rr + 2]; }
        y[3 * rr] = a0; y[3 * rr + 1] = a1; y[3 * rr + 2] = a2;
        if (partials) rowDot[i][lr] = row3<double>(d0, d1, d2, a0, a1, a2);
    };
    // short-row chunks first, from the slots loaded above; the row-loop chunks after, so their
    // loop does not hold the slot blocks live (uniform branches over the workgroup)
#pragma unroll
    for (int i = 0; i < R; i++) {
        if (partials && lane == 0) rowDot[i][lr] = 0.0;
        if (!act[i] || lng[i] || r[i] >= r1[i]) continue;
        T a0, a1, a2;
        combine16<LS>(acc[i], a0, a1, a2);
        // the row's own direction: its diagonal, the first slot of lane 0
        if (lane == 0) finish_row(i, a0, a1, a2, b[i][0].x);
    }
#pragma unroll
    for (int i = 0; i < R; i++) {
        if (!act[i] || !lng[i] || r[i] >= r1[i]) continue;
        T a0 = 0, a1 = 0, a2 = 0, px[3] = {0, 0, 0};
        row_spmv<T, FUSE, LS>(m, vals, x, r[i], lane, a0, a1, a2, z, beta[i], px);   // item 0: lane 0's px
        if (lane == 0) finish_row(i, a0, a1, a2, px);
    }
    if (partials) {
        __syncthreads();
        const int w = (int)threadIdx.x >> 6;       // wave i sums chunk i's 64 rows
        if (w < R) {
            const double s = wave_sum(rowDot[w][threadIdx.x & 63]);
            if ((threadIdx.x & 63) == 0 && chunk[w] >= 0) partials[chunk[w]] = s;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kSpmvThreads) void k_spmv(DevMesh m, const T *__restrict__ vals, const T *__restrict__ x,
                                                       T *__restrict__ y, int accumulate, const InstScalars *sc) {
    spmv_chunks<T, 1, false, kSpmvLanes>(m, vals, x, y, accumulate, nullptr, sc, nullptr, nullptr);
}

// The PCG product q = A p with the chunk partials p.q (slot 0), R chunks per workgroup and LS
// lanes per row; with FUSE the direction update p = z + beta p_old is done on the fly (x =
// p_old, p -> pNew).
template <typename T, int R, bool FUSE, int LS>
__global__ __launch_bounds__(kChunkRows * LS) void k_spmv_dot(DevMesh m, const T *__restrict__ vals,
                                                              const T *__restrict__ x, T *__restrict__ y,
                                                              double *__restrict__ partials,
                                                              const InstScalars *__restrict__ sc,
                                                              const T *__restrict__ z, T *__restrict__ pNew) {
    spmv_chunks<T, R, FUSE, LS>(m, vals, x, y, 0, partials, sc, z, pNew);
}

static inline int spmv_grid(const DevMesh &m, int R) { return 8 * ((m.perRun + R - 1) / R); }

// FP32 slot values from the FP32 upper values.  Within a run position the slots are
// slot-major (position s * 64 + r holds slot s of row r), the order the product's waves read
// them in.
__global__ __launch_bounds__(256) void k_fill_slot_values(DevMesh m, const float *__restrict__ vals,
                                                          float *__restrict__ sv) {
    // one run position per workgroup: the records are read in their row-major order (a row's
    // 16 slots by 16 consecutive threads: coalesced records, contiguous upper blocks), the
    // blocks go to LDS at their slot-major positions (row stride 65: no bank conflicts), and
    // the slot-major parts are written in position order
    constexpr int kStride = kChunkRows + 1;
    __shared__ float lds[9][kSpmvLanes * kStride];
    const int64_t p = blockIdx.x;
    const int t = (int)threadIdx.x;
#pragma unroll
    for (int j = 0; j < kSpmvThreads / 256; j++) {
        const int rec = t + 256 * j, r = rec / kSpmvLanes, sl = rec % kSpmvLanes;
        const int64_t src = p * kSpmvThreads + rec;
        const int cf = m.slotCol[src];
        float e[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (cf >= 0) {
            const float *A = vals + m.slotOff[src];
            if (cf & kSlotDiag) {
                const float d[9] = {A[0], A[1], A[2], A[1], A[3], A[4], A[2], A[4], A[5]};
#pragma unroll
                for (int q = 0; q < 9; q++) e[q] = d[q];
            } else if (cf & kSlotTrans) {
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int k = 0; k < 3; k++) e[3 * i + k] = A[3 * k + i];
            } else {
#pragma unroll
                for (int q = 0; q < 9; q++) e[q] = A[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 9; q++) lds[q][sl * kStride + r] = e[q];
    }
    __syncthreads();
    float *b = sv + p * kSlotValFloats;
#pragma unroll
    for (int j = 0; j < kSpmvThreads / 256; j++) {
        const int pi = t + 256 * j, l = (pi / kChunkRows) * kStride + pi % kChunkRows;
        *(float4 *)(b + 4 * pi) = make_float4(lds[0][l], lds[1][l], lds[2][l], lds[3][l]);
        *(float4 *)(b + 4 * kSpmvThreads + 4 * pi) = make_float4(lds[4][l], lds[5][l], lds[6][l], lds[7][l]);
        b[8 * kSpmvThreads + pi] = lds[8][l];
    }
}

void launch_fill_slot_values(const DevMesh &m, hipStream_t s, const float *vals, float *slotVals) {
    if (m.nChunks <= 0) return;
    hipLaunchKernelGGL(k_fill_slot_values, dim3((unsigned)m.nChunks), dim3(256), 0, s, m, vals, slotVals);
    OD_HIP(hipGetLastError());
}

// The FP32 PCG product on slot values (the fused direction update as in spmv_chunks): one
// chunk per workgroup, K slots per thread.  Slots are slot-major, so the wave w of 64 threads
// takes slot positions w + (16 / K) j (j < K) of the chunk's 64 rows, one row per lane: each
// wave load reads consecutive memory (1 KB per float4 part), and each gather reads the
// same-direction neighbours of 64 consecutive rows.  The row sum is the 16-lane tree of the
// other product shapes: its first levels (pairs at distance 8, then 4) inside a thread where
// its slots allow, the rest through LDS in the same pairing, so every shape gives the same
// bits.
template <int K>
__global__ __launch_bounds__(kSpmvThreads / K) void k_spmv_sv(DevMesh m, const float *__restrict__ sv,
                                                              const float *__restrict__ vals,
                                                              const float *__restrict__ x, float *__restrict__ y,
                                                              double *__restrict__ partials,
                                                              const InstScalars *__restrict__ sc,
                                                              const float *__restrict__ z, float *__restrict__ pNew) {
    constexpr int T = kSpmvThreads / K, W = kSpmvLanes / K;   // threads, waves (slot groups)
    const bool single = m.nInst == 1;
    if (single && sc[0].done) return;
    const int pos = run_pos(m, (int)(blockIdx.x >> 3)), p = pos >= 0 ? pos : 0;
    const int t = (int)threadIdx.x, r = t % kChunkRows, w = t / kChunkRows;
    const int4 info = m.runInfo[p];
    const int inst = info.w & kRunInstMask;
    // several instances: a chunk of a converged instance loads nothing (its partial is zero)
    if (!single && pos >= 0 && sc[inst].done) {
        if (partials && t == 0) partials[info.x] = 0.0;
        return;
    }
    const int *__restrict__ sCol = m.svCol + (size_t)p * kSpmvThreads;
    const float *__restrict__ vb = sv + (size_t)p * kSlotValFloats;
    int cf[K];
    BlockX<float> b[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int s = t + k * T;                 // slot w + W k of row r
        cf[k] = sCol[s];
        const float4 a0 = *(const float4 *)(vb + 4 * s), a1 = *(const float4 *)(vb + 4 * kSpmvThreads + 4 * s);
        b[k].e[0] = a0.x; b[k].e[1] = a0.y; b[k].e[2] = a0.z; b[k].e[3] = a0.w;
        b[k].e[4] = a1.x; b[k].e[5] = a1.y; b[k].e[6] = a1.z; b[k].e[7] = a1.w;
        b[k].e[8] = vb[8 * kSpmvThreads + s];
    }
    const bool act = pos >= 0;
    const bool lng = (info.w & kRunLong) != 0, use = act && !lng;
    const float beta = (float)sc[single ? 0 : inst].beta;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const bool ok = use && cf[k] >= 0;
        if (!ok) cf[k] = -1;
        const int c3 = 3 * (ok ? (cf[k] & kSlotColMask) : 0);
        b[k].x[0] = x[c3]; b[k].x[1] = x[c3 + 1]; b[k].x[2] = x[c3 + 2];
        dir_update<float>(z + c3, beta, b[k]);
    }
    __shared__ double rowDot[kChunkRows];
    __shared__ float red[W][kChunkRows][3];
    if (partials && t < kChunkRows) rowDot[t] = 0.0;
    auto finish_row = [&](int lr, float a0, float a1, float a2, const float *px) {
        const int rr = info.y + lr;
        pNew[3 * rr] = px[0]; pNew[3 * rr + 1] = px[1]; pNew[3 * rr + 2] = px[2];
        y[3 * rr] = a0; y[3 * rr + 1] = a1; y[3 * rr + 2] = a2;
        if (partials) rowDot[lr] = row3<double>(px[0], px[1], px[2], a0, a1, a2);
    };
    if (!lng) {
        float c[K][3];
#pragma unroll
        for (int k = 0; k < K; k++) {
            c[k][0] = c[k][1] = c[k][2] = 0;
            block_mv<float>(b[k], false, false, c[k][0], c[k][1], c[k][2]);
            if (cf[k] < 0) c[k][0] = c[k][1] = c[k][2] = 0;
        }
        // the tree levels inside the thread: slots (w, w + 8) [, (w + 4, w + 12), then the pair]
        float l[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if constexpr (K == 4) l[q] = (c[0][q] + c[2][q]) + (c[1][q] + c[3][q]);
            else if constexpr (K == 2) l[q] = c[0][q] + c[1][q];
            else l[q] = c[0][q];
        }
        red[w][r][0] = l[0]; red[w][r][1] = l[1]; red[w][r][2] = l[2];
        __syncthreads();
        if (w == 0) {
            float v[W][3];
#pragma unroll
            for (int i = 0; i < W; i++) { v[i][0] = red[i][r][0]; v[i][1] = red[i][r][1]; v[i][2] = red[i][r][2]; }
#pragma unroll
            for (int n = W; n > 1; n /= 2)
#pragma unroll
                for (int i = 0; i < n / 2; i++) { v[i][0] += v[i + n / 2][0]; v[i][1] += v[i + n / 2][1]; v[i][2] += v[i + n / 2][2]; }
            // the row's own direction: its diagonal, slot 0, this thread's first slot
            if (use && info.y + r < info.z) finish_row(r, v[0][0], v[0][1], v[0][2], b[0].x);
        }
    } else if (act) {
        // rows longer than 16 blocks: the row loop, 16 lanes per row (after the zeroing of
        // rowDot, which other threads did)
        __syncthreads();
        const int v16 = t % kSpmvLanes;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int lr = (t + k * T) / kSpmvLanes;
            if (info.y + lr >= info.z) continue;
            float a0 = 0, a1 = 0, a2 = 0, px[3] = {0, 0, 0};
            row_spmv<float, true, kSpmvLanes>(m, vals, x, info.y + lr, v16, a0, a1, a2, z, beta, px);
            if (v16 == 0) finish_row(lr, a0, a1, a2, px);
        }
    }
    if (partials) {
        __syncthreads();
        if (t < 64) {
            const double s = wave_sum(rowDot[t]);
            if (t == 0 && pos >= 0) partials[info.x] = s;
        }
    }
}

// ---- grouped symmetric product ---------------------------------------------------------
// Part-major storage of n positions: FP64 entries in double2 planes (0-1, 2-3, ...), FP32 in
// float4 planes (0-3, 4-7); the ninth entry of a block in its own plane, the diagonal's last
// pair (FP32) in a float2 plane.  A wave's load of one plane over 64 consecutive positions
// reads consecutive memory.
__device__ __forceinline__ void pm_load9(const double *R, int64_t n, int64_t p, double (&e)[9]) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double2 v = *(const double2 *)(R + 2 * q * n + 2 * p);
        e[2 * q] = v.x; e[2 * q + 1] = v.y;
    }
    e[8] = R[8 * n + p];
}
__device__ __forceinline__ void pm_load9(const float *R, int64_t n, int64_t p, float (&e)[9]) {
    const float4 a = *(const float4 *)(R + 4 * p), b = *(const float4 *)(R + 4 * n + 4 * p);
    e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
    e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
    e[8] = R[8 * n + p];
}
__device__ __forceinline__ void pm_store9(double *R, int64_t n, int64_t p, const double (&e)[9]) {
#pragma unroll
    for (int q = 0; q < 4; q++) *(double2 *)(R + 2 * q * n + 2 * p) = make_double2(e[2 * q], e[2 * q + 1]);
    R[8 * n + p] = e[8];
}
__device__ __forceinline__ void pm_store9(float *R, int64_t n, int64_t p, const float (&e)[9]) {
    *(float4 *)(R + 4 * p) = make_float4(e[0], e[1], e[2], e[3]);
    *(float4 *)(R + 4 * n + 4 * p) = make_float4(e[4], e[5], e[6], e[7]);
    R[8 * n + p] = e[8];
}
__device__ __forceinline__ void pm_load6(const double *R, int64_t n, int64_t p, double (&d)[6]) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double2 v = *(const double2 *)(R + 2 * q * n + 2 * p);
        d[2 * q] = v.x; d[2 * q + 1] = v.y;
    }
}
__device__ __forceinline__ void pm_load6(const float *R, int64_t n, int64_t p, float (&d)[6]) {
    const float4 a = *(const float4 *)(R + 4 * p);
    const float2 b = *(const float2 *)(R + 4 * n + 2 * p);
    d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = b.x; d[5] = b.y;
}
__device__ __forceinline__ void pm_store6(double *R, int64_t n, int64_t p, const double (&d)[6]) {
#pragma unroll
    for (int q = 0; q < 3; q++) *(double2 *)(R + 2 * q * n + 2 * p) = make_double2(d[2 * q], d[2 * q + 1]);
}
__device__ __forceinline__ void pm_store6(float *R, int64_t n, int64_t p, const float (&d)[6]) {
    *(float4 *)(R + 4 * p) = make_float4(d[0], d[1], d[2], d[3]);
    *(float2 *)(R + 4 * n + 2 * p) = make_float2(d[4], d[5]);
}

// The planes of one part-major region of n blocks as uniform base pointers, so a lane's load
// is a uniform base plus its 32-bit position.
template <typename T>
struct Planes9;
template <>
struct Planes9<double> {
    const double2 *a, *b, *c, *d;
    const double *e;
    __device__ Planes9(const double *R, int64_t n)
        : a((const double2 *)R), b((const double2 *)(R + 2 * n)), c((const double2 *)(R + 4 * n)),
          d((const double2 *)(R + 6 * n)), e(R + 8 * n) {}
    __device__ __forceinline__ void load(int p, double (&v)[9]) const {
        const double2 x0 = a[p], x1 = b[p], x2 = c[p], x3 = d[p];
        v[0] = x0.x; v[1] = x0.y; v[2] = x1.x; v[3] = x1.y; v[4] = x2.x; v[5] = x2.y; v[6] = x3.x; v[7] = x3.y;
        v[8] = e[p];
    }
};
template <>
struct Planes9<float> {
    const float4 *a, *b;
    const float *e;
    __device__ Planes9(const float *R, int64_t n) : a((const float4 *)R), b((const float4 *)(R + 4 * n)), e(R + 8 * n) {}
    __device__ __forceinline__ void load(int p, float (&v)[9]) const {
        const float4 x0 = a[p], x1 = b[p];
        v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w; v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
        v[8] = e[p];
    }
};

// 0 + the product of one block row with c, rounded as block_mv rounds a slot's first term
template <typename T>
__device__ __forceinline__ T slot_term(T e0, T e1, T e2, const T (&c)[3]) {
    T a = 0;
    acc_add(a, row3(e0, e1, e2, c[0], c[1], c[2]));
    return a;
}

// The grouped values from the canonical ones: one thread per position (upper stream, crossing
// list, then the diagonals, 64 per chunk; padding positions zero).
template <typename T>
__global__ __launch_bounds__(256) void k_fill_grp(DevMesh m, const T *__restrict__ vals, T *__restrict__ gv) {
    const int64_t Nu = m.grpNu, Nx = m.grpNx, Nd = (int64_t)m.nChunks * kChunkRows;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < Nu + Nx) {
        const bool up = p < Nu;
        const int64_t q = up ? p : p - Nu;
        const int src = up ? m.grpUpSrc[q] : m.grpXSrc[q];
        T e[9];
#pragma unroll
        for (int i = 0; i < 9; i++) e[i] = src >= 0 ? vals[(int64_t)src + i] : (T)0;
        pm_store9(up ? gv : gv + 9 * Nu, up ? Nu : Nx, q, e);
    } else if (p < Nu + Nx + Nd) {
        const int64_t q = p - Nu - Nx;
        const int c = (int)(q / kChunkRows), r = m.chunkRow0[c] + (int)(q % kChunkRows);
        T d[6];
        const bool in = r < m.chunkRow1[c];
        const int64_t src = in ? m.valOff[r] : 0;
#pragma unroll
        for (int i = 0; i < 6; i++) d[i] = in ? vals[src + i] : (T)0;
        pm_store6(gv + 9 * (Nu + Nx), Nd, q, d);
    }
}

template <typename T>
void launch_fill_grp(const DevMesh &m, hipStream_t s, const T *vals, T *grpVals) {
    const int64_t n = m.grpNu + m.grpNx + (int64_t)m.nChunks * kChunkRows;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_grp<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, vals, grpVals);
    OD_HIP(hipGetLastError());
}
template void launch_fill_grp<double>(const DevMesh &, hipStream_t, const double *, double *);
template void launch_fill_grp<float>(const DevMesh &, hipStream_t, const float *, float *);

size_t lds_max() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 65536;
    return (size_t)v;
}

// y = A x (or y += A x) over one group per workgroup, K of a row's 16 slots per thread and
// S = 16 / K threads per row: wave v of the 4 S waves serves chunk v / S of the group (one row
// per lane) with sub-index u = v mod S, and takes the row's slots u, u + S, ... (the slot-value
// kernel's assignment).
//   1. Every thread loads its slots' upper-stream entries (slot s >= 1 is stream position
//      s - 1 of the row's chunk; a row's upper slots come first, so a position holding no
//      block of this row, meta -1, is a lower or padding slot), its row's x (with FUSE the new
//      direction z + beta p_old; sub-index 0 also writes it to pNew and to LDS), and its share
//      of the group's halo (the x of the outside rows its blocks multiply) into LDS.  Loads
//      are unconditional (clamped positions, results dropped) so a thread's loads are in
//      flight together.
//   2. The upper terms A_ij x_j (x_j from LDS: the group's rows or its halo); for an internal
//      block the column row's term A_ij^T x_i goes to that row's LDS deposit, and the
//      crossing list (its owners' x from the halo) fills the other deposits.
//   3. The lower slots read their deposits; each thread sums its K slots in the 16-lane tree
//      order, the S threads of a row combine through LDS in the same order, so the bits are
//      those of the chunk-descriptor forms, and each chunk's dot partial is one wave's sum.
template <typename T, int K>
constexpr int grp_waves_per_eu() { return K == 4 ? (sizeof(T) == 8 ? 6 : 8) : (sizeof(T) == 8 ? 5 : 6); }
template <typename T, bool FUSE, int K>
__global__ __launch_bounds__(kGrpThreads * (16 / K)) __attribute__((amdgpu_waves_per_eu(grp_waves_per_eu<T, K>())))
void k_spmv_grp(DevMesh m, const T *__restrict__ gv,
                                                                     const T *__restrict__ x, T *__restrict__ y,
                                                                     int accumulate, double *__restrict__ partials,
                                                                     const InstScalars *__restrict__ sc,
                                                                     const T *__restrict__ z, T *__restrict__ pNew) {
    constexpr int S = 16 / K, NT = kGrpThreads * S;
    constexpr int KB = sizeof(T) == 8 ? 2 : 4;   // slots whose loads are in flight together
    extern __shared__ __align__(16) unsigned char grpLds[];
    T *xs = reinterpret_cast<T *>(grpLds);   // 3 per local row
    T *hx = xs + 3 * kGrpThreads;             // 3 per halo row
    T *dep = hx + 3 * m.grpMaxHalo;           // 3 per deposit; reused for the cross-thread sums
    const bool single = m.nInst == 1;
    if (single && sc && sc[0].done) return;
    const int jb = (int)(blockIdx.x >> 3), pos = (int)(blockIdx.x & 7) * m.perRunG + jb;
    if (jb >= m.perRunG || pos >= m.nGroups) return;   // whole workgroup
    const int4 gc = m.grpChunks[pos], gx = m.grpX[pos];
    const int2 gh = m.grpHalo[pos];
    const int t = (int)threadIdx.x, wv = t >> 6, lane = t & 63, ci = wv / S, u = wv % S;
    const int chunk = ci == 0 ? gc.x : (ci == 1 ? gc.y : (ci == 2 ? gc.z : gc.w));
    const int lr = ci * kChunkRows + lane;   // local row
    if (!single && sc && sc[gx.z].done) {      // whole workgroup: the group's instance is done
        if (partials && u == 0 && lane == 0 && chunk >= 0) partials[chunk] = 0.0;
        return;
    }
    const int64_t Nu = m.grpNu, Nx = m.grpNx, Nd = (int64_t)m.nChunks * kChunkRows;
    const Planes9<T> gUp(gv, Nu), gX(gv + 9 * Nu, Nx);
    const T *gD = gv + 9 * (Nu + Nx);
    int row = -1;
    int2 cu = make_int2(0, 0);
    if (chunk >= 0) {
        const int r0 = single ? chunk * kChunkRows : m.chunkRow0[chunk];
        const int r1 = single ? min(r0 + kChunkRows, m.nRows) : m.chunkRow1[chunk];
        if (r0 + lane < r1) row = r0 + lane;
        cu = m.grpChunkUp[chunk];
    }
    const T beta = FUSE ? (T)sc[single ? 0 : gx.z].beta : (T)0;
    // 1. loads
    const int rr = row >= 0 ? row : 0;
    const int info = m.grpRowInfo[rr];
    T px[3] = {x[3 * rr], x[3 * rr + 1], x[3 * rr + 2]};
    if (FUSE) {
        px[0] = fma(beta, px[0], z[3 * rr]);
        px[1] = fma(beta, px[1], z[3 * rr + 1]);
        px[2] = fma(beta, px[2], z[3 * rr + 2]);
    }
    // v[i] (i < K / 2) holds the thread's slot i and, once slot i + K / 2 (8 slots further) is
    // one of the row's own upper slots, their sum: the first level of the row's tree, added as
    // the terms come (a row's upper slots come first, so slot i is then its own term too)
    constexpr int H = K / 2;
    int meta[K];
    T v[H][3];
#pragma unroll
    for (int i = 0; i < H; i++) v[i][0] = v[i][1] = v[i][2] = 0;
    if (u == 0) {   // slot 0: the diagonal
        T d[6];
        pm_load6(gD, Nd, (int64_t)(chunk >= 0 ? chunk : 0) * kChunkRows + lane, d);
        v[0][0] = slot_term(d[0], d[1], d[2], px);
        v[0][1] = slot_term(d[1], d[3], d[4], px);
        v[0][2] = slot_term(d[2], d[4], d[5], px);
    }
    auto load_batch = [&](int i0, T (&e)[KB][9]) {
#pragma unroll
        for (int i = 0; i < KB; i++) {
            const int sl = u + S * (i0 + i), k = sl - 1;
            const bool in = row >= 0 && sl >= 1 && k < cu.y;
            const int p = in ? cu.x + k * kChunkRows + lane : 0;
            meta[i0 + i] = m.grpUpMeta[p];
            gUp.load(p, e[i]);
            if (!in) meta[i0 + i] = -1;
        }
    };
    // column entries and terms of a loaded batch (after the LDS x is written)
    auto use_batch = [&](int i0, const T (&e)[KB][9]) {
#pragma unroll
        for (int i = 0; i < KB; i++) {
            const int mt = meta[i0 + i];
            if (mt == -1) continue;   // not an upper slot of this row (the diagonal, a lower slot, padding)
            // the column's x: a halo row's, or (internal block) a row of the group
            const T *cx = mt >= 0 ? hx + 3 * mt : xs + 3 * ((mt >> 14) & 0xFF);
            const T c[3] = {cx[0], cx[1], cx[2]};
            const T o0 = slot_term(e[i][0], e[i][1], e[i][2], c);
            const T o1 = slot_term(e[i][3], e[i][4], e[i][5], c);
            const T o2 = slot_term(e[i][6], e[i][7], e[i][8], c);
            if (i0 + i < H) {
                v[i0 + i][0] = o0; v[i0 + i][1] = o1; v[i0 + i][2] = o2;
            } else {
                v[i0 + i - H][0] += o0; v[i0 + i - H][1] += o1; v[i0 + i - H][2] += o2;
            }
            if (mt < 0 && mt != -1) {
                T *o = dep + 3 * (mt & 0x3FFF);
                o[0] = slot_term(e[i][0], e[i][3], e[i][6], px);
                o[1] = slot_term(e[i][1], e[i][4], e[i][7], px);
                o[2] = slot_term(e[i][2], e[i][5], e[i][8], px);
            }
        }
    };
    T e0[KB][9];
    load_batch(0, e0);
    // the crossing list: at most two items per thread in flight
    constexpr int XB = 1;
    int2 xm[XB];
    T xe[XB][9];
    auto load_x = [&](int q0) {
#pragma unroll
        for (int i = 0; i < XB; i++) {
            const int q = q0 + i * NT;
            const int p = gx.x + (q < gx.y ? q : 0);
            xm[i] = m.grpXMeta[p];
            gX.load(p, xe[i]);
            if (q >= gx.y) xm[i].y = -1;
        }
    };
    auto use_x = [&]() {
#pragma unroll
        for (int i = 0; i < XB; i++) {
            if (xm[i].y < 0) continue;
            const T c[3] = {hx[3 * xm[i].x], hx[3 * xm[i].x + 1], hx[3 * xm[i].x + 2]};
            T *o = dep + 3 * xm[i].y;
            o[0] = slot_term(xe[i][0], xe[i][3], xe[i][6], c);
            o[1] = slot_term(xe[i][1], xe[i][4], xe[i][7], c);
            o[2] = slot_term(xe[i][2], xe[i][5], xe[i][8], c);
        }
    };
    if (u == 0) {
        xs[3 * lr] = px[0]; xs[3 * lr + 1] = px[1]; xs[3 * lr + 2] = px[2];
        if (FUSE && row >= 0) { pNew[3 * row] = px[0]; pNew[3 * row + 1] = px[1]; pNew[3 * row + 2] = px[2]; }
    }
    // the halo's x (with FUSE its direction z + beta p_old, formed as the rows' own)
    for (int h = t; h < gh.y; h += NT) {
        const int hr = m.grpHaloRows[gh.x + h];
        T c[3] = {x[3 * hr], x[3 * hr + 1], x[3 * hr + 2]};
        if (FUSE) {
            c[0] = fma(beta, c[0], z[3 * hr]);
            c[1] = fma(beta, c[1], z[3 * hr + 1]);
            c[2] = fma(beta, c[2], z[3 * hr + 2]);
        }
        hx[3 * h] = c[0]; hx[3 * h + 1] = c[1]; hx[3 * h + 2] = c[2];
    }
    __syncthreads();
    // 2. terms and deposits
    use_batch(0, e0);
#pragma unroll
    for (int i0 = KB; i0 < K; i0 += KB) {
        T e1[KB][9];
        load_batch(i0, e1);
        use_batch(i0, e1);
    }
    for (int q0 = t; q0 < gx.y; q0 += XB * NT) {
        load_x(q0);
        use_x();
    }
    __syncthreads();
    // 3. slot values, the thread's tree levels, then the row's S threads through LDS
    const int nUp = info & 15, items = (info >> 4) & 31, db = info >> 9;
    auto slot_value = [&](int sl, T (&a)[3]) {   // a lower slot's deposit, or padding
        const bool has = sl < items && row >= 0;
        const int o = 3 * (has ? db + sl - nUp - 1 : 0);
        a[0] = has ? dep[o] : (T)0;
        a[1] = has ? dep[o + 1] : (T)0;
        a[2] = has ? dep[o + 2] : (T)0;
    };
#pragma unroll
    for (int i = 0; i < H; i++) {
        const int sl = u + S * i;
        if (sl != 0 && meta[i] == -1) slot_value(sl, v[i]);      // slot i is no own term
        if (meta[i + H] == -1) {                                 // slot i + H is not folded in yet
            T a[3];
            slot_value(sl + 8, a);
            v[i][0] += a[0]; v[i][1] += a[1]; v[i][2] += a[2];
        }
    }
#pragma unroll
    for (int o = H / 2; o > 0; o >>= 1)
#pragma unroll
        for (int i = 0; i < o; i++) {
            v[i][0] += v[i + o][0]; v[i][1] += v[i + o][1]; v[i][2] += v[i + o][2];
        }
    T a0 = v[0][0], a1 = v[0][1], a2 = v[0][2];
    if constexpr (S > 1) {
        __syncthreads();                       // every deposit read: the area takes the sums
        T *red = dep;                          // [S - 1][kGrpThreads][3]
        if (u > 0) {
            T *o = red + 3 * ((u - 1) * kGrpThreads + lr);
            o[0] = a0; o[1] = a1; o[2] = a2;
        }
        __syncthreads();
        if (u == 0) {
            T r[S][3];
            r[0][0] = a0; r[0][1] = a1; r[0][2] = a2;
#pragma unroll
            for (int j = 1; j < S; j++) {
                const T *o = red + 3 * ((j - 1) * kGrpThreads + lr);
                r[j][0] = o[0]; r[j][1] = o[1]; r[j][2] = o[2];
            }
#pragma unroll
            for (int o = S / 2; o > 0; o >>= 1)
#pragma unroll
                for (int j = 0; j < o; j++) {
                    r[j][0] += r[j + o][0]; r[j][1] += r[j + o][1]; r[j][2] += r[j + o][2];
                }
            a0 = r[0][0]; a1 = r[0][1]; a2 = r[0][2];
        }
    }
    if (u != 0) return;                        // no barrier below
    double rowDot = 0.0;
    if (row >= 0) {
        if (accumulate) { a0 += y[3 * row]; a1 += y[3 * row + 1]; a2 += y[3 * row + 2]; }
        y[3 * row] = a0; y[3 * row + 1] = a1; y[3 * row + 2] = a2;
        if (partials) rowDot = row3<double>(px[0], px[1], px[2], a0, a1, a2);
    }
    if (partials) {
        const double sum = wave_sum(rowDot);
        if (lane == 0 && chunk >= 0) partials[chunk] = sum;
    }
}

template <typename T>
void launch_spmv_dot(const DevMesh &m, hipStream_t s, int R, int LS, const T *vals, const T *x, T *y, double *partials,
                     const InstScalars *sc, hipEvent_t start, hipEvent_t stop, const T *z, T *pNew,
                     const T *slotVals) {
    if (LS == -1) {
        if (!slotVals) throw HipError("the grouped product needs its grouped values");
        if (!(sizeof(T) == 8 ? m.grpOk64 : m.grpOk32)) throw HipError("the grouped product is not available on this mesh");
        if ((sizeof(T) == 4) != (z != nullptr)) throw HipError("the FP32 grouped product runs with the fused direction update");
        if (z && !sc) throw HipError("the fused direction update needs the solver scalars");
        if (R != 4 && R != 8) throw HipError("the grouped product runs 4 or 8 slots per thread");
        const size_t lds = grp_lds_bytes<T>(m, R);
        const dim3 grid(8 * (unsigned)m.perRunG);
        auto go = [&](auto kernel, int threads) {
            // LDS beyond 64 KB (gfx950 gives a workgroup up to the CU's 160 KB): the kernel's
            // dynamic limit is raised to it once
            static bool raised = false;
            if (lds > 65536 && !raised) {
                OD_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max()));
                raised = true;
            }
            if (start)
                hipExtLaunchKernelGGL(kernel, grid, dim3(threads), lds, s, start, stop, 0, m, slotVals, x, y, 0,
                                      partials, sc, z, pNew);
            else
                hipLaunchKernelGGL(kernel, grid, dim3(threads), lds, s, m, slotVals, x, y, 0, partials, sc, z, pNew);
        };
        constexpr bool F = sizeof(T) == 4;
        if (R == 4) go(k_spmv_grp<T, F, 4>, kGrpThreads * 4);
        else go(k_spmv_grp<T, F, 8>, kGrpThreads * 2);
        OD_HIP(hipGetLastError());
        return;
    }
    // With events, the launch records them from the dispatch's own begin/end timestamps,
    // so their elapsed time is the kernel's execution alone (measurement mode).
    auto go = [&](auto kernel, int threads) {
        if (start)
            hipExtLaunchKernelGGL(kernel, dim3(spmv_grid(m, R)), dim3(threads), 0, s, start, stop, 0, m, vals, x, y,
                                  partials, sc, z, pNew);
        else
            hipLaunchKernelGGL(kernel, dim3(spmv_grid(m, R)), dim3(threads), 0, s, m, vals, x, y, partials, sc, z, pNew);
    };
    auto shape = [&](auto fuse) {
        constexpr bool F = decltype(fuse)::value;
        if (LS == 4) {
            if (R == 4) go(k_spmv_dot<T, 4, F, 4>, kChunkRows * 4);
            else if (R == 2) go(k_spmv_dot<T, 2, F, 4>, kChunkRows * 4);
            else go(k_spmv_dot<T, 1, F, 4>, kChunkRows * 4);
        } else if (LS == 8) {
            if (R == 4) go(k_spmv_dot<T, 4, F, 8>, kChunkRows * 8);
            else if (R == 2) go(k_spmv_dot<T, 2, F, 8>, kChunkRows * 8);
            else go(k_spmv_dot<T, 1, F, 8>, kChunkRows * 8);
        } else {
            if (R == 4) go(k_spmv_dot<T, 4, F, 16>, kChunkRows * 16);
            else if (R == 2) go(k_spmv_dot<T, 2, F, 16>, kChunkRows * 16);
            else go(k_spmv_dot<T, 1, F, 16>, kChunkRows * 16);
        }
    };
    if constexpr (sizeof(T) == 4) {          // the fused direction update is an FP32 product
        if (!z) throw HipError("the FP32 product runs with the fused direction update");
        if (LS == 0) {
            if (!slotVals || !sc) throw HipError("the slot-value product needs the slot values and scalars");
            auto sv = [&](auto kernel, int threads) {
                if (start)
                    hipExtLaunchKernelGGL(kernel, dim3(spmv_grid(m, 1)), dim3(threads), 0, s, start, stop, 0, m,
                                          slotVals, vals, x, y, partials, sc, z, pNew);
                else
                    hipLaunchKernelGGL(kernel, dim3(spmv_grid(m, 1)), dim3(threads), 0, s, m, slotVals, vals, x, y,
                                       partials, sc, z, pNew);
            };
            if (R == 4) sv(k_spmv_sv<4>, kSpmvThreads / 4);
            else if (R == 2) sv(k_spmv_sv<2>, kSpmvThreads / 2);
            else sv(k_spmv_sv<1>, kSpmvThreads);
        } else {
            shape(std::true_type{});
        }
    } else {
        if (z) throw HipError("the fused direction update is an FP32 product");
        if (LS == 0) throw HipError("the slot-value product is an FP32 product");
        shape(std::false_type{});
    }
    OD_HIP(hipGetLastError());
}
template void launch_spmv_dot<double>(const DevMesh &, hipStream_t, int, int, const double *, const double *, double *,
                                      double *, const InstScalars *, hipEvent_t, hipEvent_t, const double *, double *,
                                      const double *);
template void launch_spmv_dot<float>(const DevMesh &, hipStream_t, int, int, const float *, const float *, float *,
                                     double *, const InstScalars *, hipEvent_t, hipEvent_t, const float *, float *,
                                     const float *);

template <typename T>
void launch_spmv(const DevMesh &m, hipStream_t s, const T *vals, const T *x, T *y, int accumulate,
                 const InstScalars *sc, const T *grpVals) {
    // an FP64 product y = A x runs the autotuned FP64 shape (same bits as every shape)
    if constexpr (sizeof(T) == 8) {
        if (!accumulate) {
            if (m.spmvL64 == -1 && grpVals)
                launch_spmv_dot<T>(m, s, m.spmvR64, -1, vals, x, y, nullptr, sc, nullptr, nullptr, nullptr, nullptr,
                                   grpVals);
            else if (m.spmvL64 == -1)
                launch_spmv_dot<T>(m, s, m.spmvR64d, m.spmvL64d, vals, x, y, nullptr, sc);
            else
                launch_spmv_dot<T>(m, s, m.spmvR64, m.spmvL64, vals, x, y, nullptr, sc);
            return;
        }
    }
    hipLaunchKernelGGL(k_spmv<T>, dim3(spmv_grid(m, 1)), dim3(kSpmvThreads), 0, s, m, vals, x, y, accumulate, sc);
    OD_HIP(hipGetLastError());
}
template void launch_spmv<double>(const DevMesh &, hipStream_t, const double *, const double *, double *, int,
                                  const InstScalars *, const double *);
template void launch_spmv<float>(const DevMesh &, hipStream_t, const float *, const float *, float *, int,
                                 const InstScalars *, const float *);

template <typename T>
void pcg_alpha(const DevMesh &m, hipStream_t s, const T *vals, const T *pOld, const T *z, T *p, T *q, PcgWork &w,
               hipEvent_t start, hipEvent_t stop) {
    int R = sizeof(T) == 8 ? m.spmvR64 : m.spmvR32, L = sizeof(T) == 8 ? m.spmvL64 : m.spmvL32;
    if (sizeof(T) == 8 && L == -1 && !w.slotVals) {   // no grouped values (the allocation-free solve)
        R = m.spmvR64d;
        L = m.spmvL64d;
    }
    launch_spmv_dot<T>(m, s, R, L, vals, z ? pOld : p, q, w.partials, w.sc, start, stop, z, z ? p : nullptr,
                       L <= 0 ? (const T *)w.slotVals : nullptr);
}

void pcg_alpha_finalize(const DevMesh &m, hipStream_t s, PcgWork &w) {
    launch_finalize(m, s, PH_ALPHA, w, 0.0, 0, nullptr);
}

// =====================================================================================
// PCG vector kernels: one workgroup per chunk, one DOF per thread
// =====================================================================================
constexpr int kVecThreads = 3 * kChunkRows;

// A mesh of one instance has chunk c on rows [64 c, min(64 c + 64, nRows)): no chunk record
// is loaded before the vectors.
#define CHUNK_PROLOGUE                                                                        \
    const int chunk = blockIdx.x;                                                             \
    const bool singleInst = m.nInst == 1;                                                     \
    const int inst = singleInst ? 0 : m.chunkInst[chunk];                                     \
    const int row0 = singleInst ? chunk * kChunkRows : m.chunkRow0[chunk];                    \
    const int row1 = singleInst ? min(row0 + kChunkRows, m.nRows) : m.chunkRow1[chunk];       \
    const int i = 3 * row0 + (int)threadIdx.x;                                                \
    const bool in = i < 3 * row1;                                                             \
    (void)inst;

// r = b - q ; slots [1] r.r, [2] b.b
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_init_res(DevMesh m, const T *b, const T *q, T *r, double *partials) {
    CHUNK_PROLOGUE
    double v[2] = {0, 0};
    if (in) {
        T ri = b[i] - q[i];
        r[i] = ri;
        v[0] = (double)ri * (double)ri;
        v[1] = (double)b[i] * (double)b[i];
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) { partials[m.nChunks + chunk] = v[0]; partials[2 * m.nChunks + chunk] = v[1]; }
}

template <typename T>
void pcg_init_residual(const DevMesh &m, hipStream_t s, const T *b, const T *q, T *r, PcgWork &w) {
    hipLaunchKernelGGL(k_init_res<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, b, q, r, w.partials);
    OD_HIP(hipGetLastError());
}

// z = invDiag * r (no reduction); invDiag == nullptr: identity
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_jacobi(DevMesh m, const T *invDiag, const T *r, T *z) {
    CHUNK_PROLOGUE
    if (in) z[i] = invDiag ? invDiag[i] * r[i] : r[i];
}

template <typename T>
void pcg_jacobi(const DevMesh &m, hipStream_t s, const T *invDiag, const T *r, T *z, PcgWork &w) {
    (void)w;
    hipLaunchKernelGGL(k_jacobi<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, invDiag, r, z);
    OD_HIP(hipGetLastError());
}

// p = z ; slot [0] r.z ; then finalise: delta, eps, done
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_init_rz(DevMesh m, const T *r, const T *z, T *p, double *partials) {
    CHUNK_PROLOGUE
    double v[1] = {0};
    if (in) {
        p[i] = z[i];
        v[0] = (double)r[i] * (double)z[i];
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[chunk] = v[0];
}

template <typename T>
void pcg_dot_rz_init(const DevMesh &m, hipStream_t s, const T *r, const T *z, T *p, PcgWork &w, double tol, int maxIt,
                     const uint8_t *preDone) {
    hipLaunchKernelGGL(k_init_rz<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, r, z, p, w.partials);
    OD_HIP(hipGetLastError());
    launch_finalize(m, s, PH_INIT, w, tol, maxIt, preDone);
}

// x += alpha p ; r -= alpha q ; slot [1] r.r ; fused Jacobi: z = D^-1 r and slot [0] r.z (+ finalise)
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_update_xr(DevMesh m, const T *p, const T *q, T *x, T *r, const T *invDiag,
                                                           T *z, int fuseJacobi, double *partials,
                                                           const InstScalars *sc) {
    CHUNK_PROLOGUE
    double v[2] = {0, 0};
    if (in && !sc[inst].done) {
        const T a = (T)sc[inst].alpha;
        T xi = x[i] + a * p[i];
        T ri = r[i] - a * q[i];
        x[i] = xi;
        r[i] = ri;
        v[1] = (double)ri * (double)ri;
        if (fuseJacobi) {
            T zi = invDiag ? invDiag[i] * ri : ri;
            z[i] = zi;
            v[0] = (double)ri * (double)zi;
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        partials[m.nChunks + chunk] = v[1];
        if (fuseJacobi) partials[chunk] = v[0];
    }
}

template <typename T>
void pcg_update_xr(const DevMesh &m, hipStream_t s, const T *p, const T *q, T *x, T *r, const T *invDiag, T *z,
                   PcgWork &w) {
    hipLaunchKernelGGL(k_update_xr<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, p, q, x, r, invDiag, z, z ? 1 : 0,
                       w.partials, w.sc);
    OD_HIP(hipGetLastError());
    if (z) launch_finalize(m, s, PH_ITER, w, 0.0, 0, nullptr);
}

// slot [0] r.z ; finalise beta, delta, iteration count, done
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_dot_rz(DevMesh m, const T *r, const T *z, double *partials,
                                                        const InstScalars *sc) {
    CHUNK_PROLOGUE
    double v[1] = {0};
    if (in && !sc[inst].done) v[0] = (double)r[i] * (double)z[i];
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[chunk] = v[0];
}

template <typename T>
void pcg_dot_rz(const DevMesh &m, hipStream_t s, const T *r, const T *z, PcgWork &w) {
    hipLaunchKernelGGL(k_dot_rz<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, r, z, w.partials, w.sc);
    OD_HIP(hipGetLastError());
    launch_finalize(m, s, PH_ITER, w, 0.0, 0, nullptr);
}

// p = z + beta p (skipped for instances that are done): the direction update of the FP64
// solve, whose products run without the fused update (the descriptor shapes would need more
// registers than two workgroups per CU allow; the grouped product measured slower with it
// than this separate pass)
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_update_p(DevMesh m, const T *z, T *p, const InstScalars *sc) {
    CHUNK_PROLOGUE
    if (!in || sc[inst].done) return;
    p[i] = fma((T)sc[inst].beta, p[i], z[i]);
}

template <typename T>
void pcg_update_p(const DevMesh &m, hipStream_t s, const T *z, T *p, PcgWork &w) {
    hipLaunchKernelGGL(k_update_p<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, z, p, w.sc);
    OD_HIP(hipGetLastError());
}

// inverse diagonal, 1 where the diagonal is <= 0 (jacobiPreconditioner.cpp:18-21)
template <typename T>
__global__ void k_jacobi_diag(DevMesh m, const double *vals, T *inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * m.nRows) return;
    const int r = i / 3, ax = i % 3;
    const int off = m.valOff[r] + (ax == 0 ? 0 : (ax == 1 ? 3 : 5));
    double dg = vals[off];
    inv[i] = (T)(dg > 0.0 ? 1.0 / dg : 1.0);
}

template <typename T>
void launch_jacobi_diag(const DevMesh &m, hipStream_t s, const double *vals, T *invDiag) {
    int n = 3 * m.nRows;
    hipLaunchKernelGGL(k_jacobi_diag<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m, vals, invDiag);
    OD_HIP(hipGetLastError());
}

// =====================================================================================
// IC(0), block form: U upper (canonical layout), U^T U = A on the block pattern
// =====================================================================================
__device__ __forceinline__ double pivot_fix(double d, double eps2) { return d <= 0.0 ? eps2 * d * d : d; }

// Entry q (0-8) of block b in a part-major array of nb blocks (the FP32 factor copies).
__device__ __forceinline__ int64_t pm_at(int64_t nb, int64_t b, int q) {
    return q < 4 ? 4 * b + q : (q < 8 ? 4 * nb + 4 * b + (q - 4) : 8 * nb + b);
}

// The nine entries of block b of a part-major factor copy of nb blocks: FP32 two float4 loads
// and one float, FP64 four double2 loads and one double.
__device__ __forceinline__ void load_pm(const float *__restrict__ P, int64_t nb, int64_t b, float (&e)[9]) {
    const float4 a = ((const float4 *)P)[b], c = ((const float4 *)(P + 4 * nb))[b];
    e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
    e[4] = c.x; e[5] = c.y; e[6] = c.z; e[7] = c.w;
    e[8] = P[8 * nb + b];
}
__device__ __forceinline__ void load_pm(const double *__restrict__ P, int64_t nb, int64_t b, double (&e)[9]) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double2 v = ((const double2 *)(P + 2 * q * nb))[b];
        e[2 * q] = v.x; e[2 * q + 1] = v.y;
    }
    e[8] = P[8 * nb + b];
}
// Block b of a part-major copy from FP64 entries, rounded to the copy's type.
__device__ __forceinline__ void store_pm(float *P, int64_t nb, int64_t b, const double (&X)[9]) {
    ((float4 *)P)[b] = make_float4((float)X[0], (float)X[1], (float)X[2], (float)X[3]);
    ((float4 *)(P + 4 * nb))[b] = make_float4((float)X[4], (float)X[5], (float)X[6], (float)X[7]);
    P[8 * nb + b] = (float)X[8];
}
__device__ __forceinline__ void store_pm(double *P, int64_t nb, int64_t b, const double (&X)[9]) {
#pragma unroll
    for (int q = 0; q < 4; q++) ((double2 *)(P + 2 * q * nb))[b] = make_double2(X[2 * q], X[2 * q + 1]);
    P[8 * nb + b] = X[8];
}

// One block row of U per 8-lane group: lane 0 forms and factors the diagonal block,
// lanes 1..7 the row's upper off-diagonal blocks U_IJ = U_II^-T (A_IJ - sum_K U_KI^T U_KJ),
// the terms K from the block's precomputed pair list (icPairs: no search).  Every U_KI is
// read from U itself (transposed value offsets).  The sweeps' copies (T: FP32 for the mixed
// form, FP64 for the FP64 sweeps) are written when given: Uc the diagonal blocks (8 entries
// per row) then the upper blocks at their backward sweep slots, Ulc the lower blocks at their
// forward sweep slots, both part-major.

template <typename T>
__global__ __launch_bounds__(256) void k_ic0_factor_level(DevMesh m, const int *__restrict__ rows, int count,
                                                          const double *__restrict__ A, double *__restrict__ U,
                                                          T *__restrict__ U32, T *__restrict__ Ul32, double eps2,
                                                          const uint8_t *__restrict__ skip) {
    const int wg = ((count + 31) >> 5) > 0 ? ((int)(blockIdx.x & 7) * ((((count + 31) >> 5) + 7) >> 3) + (int)(blockIdx.x >> 3)) : 0;
    const int t = wg * 32 + (threadIdx.x >> 3), lane = threadIdx.x & 7;
    if (t >= count) return;
    const int I = rows[t];
    if (skip && skip[m.rowInst[I]]) return;
    const int b0 = m.rowPtr[I], nOff = m.rowPtr[I + 1] - b0 - 1, v0 = m.valOff[I];
    const int k0 = m.tPtr[I], k1 = m.tPtr[I + 1];
    double u[6] = {0, 0, 0, 0, 0, 0};
    // sum of U_KI^T U_KI over the row's lower blocks, spread over the row's 8 lanes (lane l takes
    // blocks l, l + 8, ...) and combined in a fixed tree (pairs at distance 4, 2, 1)
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (int k = k0 + lane; k < k1; k += 8) {
        const double *W = U + m.tOff[k];
        const double w0 = W[0], w1 = W[1], w2 = W[2], w3 = W[3], w4 = W[4], w5 = W[5], w6 = W[6], w7 = W[7], w8 = W[8];
        g[0] += w0 * w0 + w3 * w3 + w6 * w6;
        g[1] += w0 * w1 + w3 * w4 + w6 * w7;
        g[2] += w0 * w2 + w3 * w5 + w6 * w8;
        g[3] += w1 * w1 + w4 * w4 + w7 * w7;
        g[4] += w1 * w2 + w4 * w5 + w7 * w8;
        g[5] += w2 * w2 + w5 * w5 + w8 * w8;
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < 6; j++) g[j] += __shfl_xor(g[j], o, 8);
    if (lane == 0) {
        const double s00 = A[v0] - g[0], s01 = A[v0 + 1] - g[1], s02 = A[v0 + 2] - g[2];
        const double s11 = A[v0 + 3] - g[3], s12 = A[v0 + 4] - g[4], s22 = A[v0 + 5] - g[5];
        // dense Cholesky of the 3x3 diagonal block, pivot rule d <= 0 -> eps^2 d^2
        u[0] = sqrt(pivot_fix(s00, eps2));
        u[1] = s01 / u[0];
        u[2] = s02 / u[0];
        u[3] = sqrt(pivot_fix(s11 - u[1] * u[1], eps2));
        u[4] = (s12 - u[1] * u[2]) / u[3];
        u[5] = sqrt(pivot_fix(s22 - u[2] * u[2] - u[4] * u[4], eps2));
#pragma unroll
        for (int j = 0; j < 6; j++) {
            U[v0 + j] = u[j];
            if (U32) U32[8 * (int64_t)I + j] = (T)u[j];
        }
    }
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = __shfl(u[j], 0, 8);
    if (lane == 0) return;
    for (int q = lane - 1; q < nOff; q += 7) {
        const int vq = v0 + 6 + 9 * q;
        double S[9];
#pragma unroll
        for (int j = 0; j < 9; j++) S[j] = A[vq + j];
        const int e1 = m.icPairPtr[b0 + 2 + q];
        for (int e = m.icPairPtr[b0 + 1 + q]; e < e1; e++) {
            const int2 pr = m.icPairs[e];
            const double *W = U + pr.x, *Z = U + pr.y;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) S[i * 3 + j] -= W[0 * 3 + i] * Z[0 * 3 + j] + W[1 * 3 + i] * Z[1 * 3 + j] + W[2 * 3 + i] * Z[2 * 3 + j];
        }
        // U_IJ = U_II^-T S (forward substitution with the lower factor U_II^T)
        double X[9];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            X[j] = S[j] / u[0];
            X[3 + j] = (S[3 + j] - u[1] * X[j]) / u[3];
            X[6 + j] = (S[6 + j] - u[2] * X[j] - u[4] * X[3 + j]) / u[5];
        }
        const int64_t bq = b0 + 1 + q, tq = m.tPos[bq];
#pragma unroll
        for (int j = 0; j < 9; j++) U[vq + j] = X[j];
        // U_IJ at slot q of row I's backward list, and as block tq - tPtr[J] of row J's forward list
        if (U32) store_pm(U32 + 8 * (int64_t)m.nRows, m.swNB, m.swBaseB[I] + (int64_t)q * kChunkRows, X);
        if (Ul32) {
            const int J = m.colIdx[bq];
            store_pm(Ul32, m.swNF, m.swBaseF[J] + (tq - m.tPtr[J]) * kChunkRows, X);
        }
    }
}

template <typename T>
void launch_ic0_factor(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, const double *A, double *U,
                       T *U32, T *Ul32, double pivotEps, const uint8_t *skip) {
    const double eps2 = pivotEps * pivotEps;
    for (size_t l = 0; l + 1 < levelPtr.size(); l++) {
        int cnt = levelPtr[l + 1] - levelPtr[l];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_ic0_factor_level<T>, dim3(8 * ((((cnt + 31) / 32) + 7) / 8)), dim3(256), 0, s, m,
                           m.levelRows + levelPtr[l], cnt, A, U, U32, Ul32, eps2, skip);
    }
    OD_HIP(hipGetLastError());
}

// XCD-aware workgroup index for level kernels: the launch has 8*ceil(nwg/8) workgroups;
// XCD x (= blockIdx mod 8) runs the contiguous run [x*per, (x+1)*per) of the level's rows.
__device__ __forceinline__ int xcd_wg(int nwg) {
    const int per = (nwg + 7) >> 3;
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}
static inline int xcd_grid(int nwg) { return 8 * ((nwg + 7) / 8); }

// IC(0) sweeps, one launch per level (rows of a level are independent), L lanes per row.
// In a colour-grouped numbering a row of the first colour has all its neighbours above it
// and a row of the last colour all below, so in those levels one side holds up to 14 blocks
// of a Kuhn grid: levels whose rows average more than 8 blocks on the swept side run 16
// lanes per row, the others 8.
constexpr int kIcThreads = 256;

// y <- U_II^-1 y for the packed upper diagonal block D = [d0 d1 d2; . d3 d4; . . d5]
template <typename T>
__device__ __forceinline__ void diag_bwd(const T *D, T &y0, T &y1, T &y2) {
    y2 = y2 / D[5];
    y1 = (y1 - D[4] * y2) / D[3];
    y0 = (y0 - D[1] * y1 - D[2] * y2) / D[0];
}

// Whether a sweep row still runs: instance inst neither converged (sc) nor skipped.  Called
// twice per kernel: first (head) with instance 0 for a mesh of one instance, from fixed
// addresses before any row data is loaded (the whole grid exits together past convergence),
// then with the row's instance for a mesh of several.
// Absent scalars or skip flags read as these (zero: live), so the two flags are loaded
// together, without a branch on either pointer.
__device__ InstScalars g_liveScalars;
__device__ uint8_t g_liveSkip[1];

__device__ __forceinline__ bool sweep_row_live(const DevMesh &m, int inst, const InstScalars *__restrict__ sc,
                                               const uint8_t *__restrict__ skip, bool head) {
    if ((m.nInst == 1) != head) return true;
    const int d = (sc ? sc : &g_liveScalars)[sc ? inst : 0].done;
    const int k = (skip ? skip : g_liveSkip)[skip ? inst : 0];
    return (d | k) == 0;
}

// A factor block and the three entries of x it multiplies, from a part-major sweep copy (slot
// b of nb).
template <typename T>
__device__ __forceinline__ void load_factor_block(const T *__restrict__ P, int64_t nb, int64_t b,
                                                  const T *__restrict__ xc, BlockX<T> &blk) {
    load_pm(P, nb, b, blk.e);
    blk.x[0] = xc[0]; blk.x[1] = xc[1]; blk.x[2] = xc[2];
}

// The packed diagonal block of row I from the diagonal part of a sweep copy (8 entries per row).
template <typename T>
__device__ __forceinline__ void load_diag(const T *__restrict__ U, int I, T (&d)[6]) {
    if constexpr (sizeof(T) == 4) {
        const float4 a = ((const float4 *)U)[2 * (int64_t)I];
        const float2 c = ((const float2 *)U)[4 * (int64_t)I + 2];
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = c.x; d[5] = c.y;
    } else {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double2 v = ((const double2 *)U)[4 * (int64_t)I + q];
            d[2 * q] = v.x; d[2 * q + 1] = v.y;
        }
    }
}

// Rows of a sweep workgroup.  Plain launches: kIcThreads / L consecutive rows of the level
// (level-relative).  Launches that also form the r.z dot (DOT = 1: one instance with levels
// as consecutive row ranges; DOT = 2: several instances, over the level's chunk list): the
// rows of one 64-row chunk that lie in the level, so the workgroup can reduce the chunk's r.z
// terms.
// xcd_wg spreads either over the XCDs.
__device__ __forceinline__ int sweep_chunks(int p0, int count) {
    return (p0 + count - 1) / kChunkRows - p0 / kChunkRows + 1;
}

// The chunk partial of r.z from the final z of the rows of chunk c in this level: the 192
// products r_i z_i of the chunk (zero outside the level) summed as k_dot_rz sums them (one
// wave per 64 entries, then the three wave sums in order).  The chunk's first launch (the
// highest level it touches: levels finish in decreasing order) stores the sum, later ones add
// theirs to it.  Every thread of the workgroup calls this.
template <int DOT>
__device__ __forceinline__ void sweep_dot(const DevMesh &m, int c, int p0, int count, int level, int lr, bool writer,
                                          const double (&v)[3], double *__restrict__ partials) {
    __shared__ double rz[3 * kChunkRows];
    __shared__ double ws[3];
    if (writer) { rz[3 * lr] = v[0]; rz[3 * lr + 1] = v[1]; rz[3 * lr + 2] = v[2]; }
    __syncthreads();
    for (int q = threadIdx.x; q < 3 * kChunkRows; q += blockDim.x) {   // whole waves
        const double s = wave_sum(rz[q]);
        if ((q & 63) == 0) ws[q >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < 3; k++) s += ws[k];
        const int cEnd = min((c + 1) * kChunkRows, m.nRows);
        bool first;
        if constexpr (DOT == 1) first = cEnd <= p0 + count;
        else first = level == m.chunkTopLevel[c];
        partials[c] = first ? s : partials[c] + s;
    }
}

// forward: U^T y = r  ->  y_I = U_II^-T (r_I - sum_{K<I} U_KI^T y_K), the U_KI from the
// lower copy (FP64: contiguous per row I; FP32: row I's forward sweep slots)
template <typename T, int L, int DOT>
__global__ __launch_bounds__(DOT ? kChunkRows * L : kIcThreads) void k_ic0_fwd(
    DevMesh m, int p0, int count, const T *__restrict__ U, const T *__restrict__ Ul, const T *__restrict__ r,
    T *__restrict__ y, const InstScalars *__restrict__ sc, const uint8_t *__restrict__ skip, int lastLevel,
    double *__restrict__ partials, int level) {
    constexpr int kRows = DOT ? kChunkRows : kIcThreads / L, G = 16 / L;
    const int lane = threadIdx.x % L, lr = threadIdx.x / L;
    int I, c = 0;
    bool live;
    if constexpr (DOT) {
        if constexpr (DOT == 1) {
            const int nwg = sweep_chunks(p0, count), wg = xcd_wg(nwg);
            if (wg >= nwg) return;              // grid padding: the whole workgroup
            c = p0 / kChunkRows + wg;
            I = c * kChunkRows + lr;
            live = I >= p0 && I < p0 + count;
        } else {                                // several instances: p0 and count name the
            const int wg = xcd_wg(count);       // level's range of the chunk list
            if (wg >= count) return;
            c = m.levelChunks[p0 + wg];
            I = m.chunkRow0[c] + lr;
            live = I < m.chunkRow1[c] && m.rowLevel[I] == level;
        }
    } else {
        const int pl = xcd_wg((count + kRows - 1) / kRows) * kRows + lr;
        if (pl >= count) return;
        I = m.levelsContiguous ? p0 + pl : m.levelRows[p0 + pl];
        live = true;
    }
    int4 ri = make_int4(0, 0, 0, 0);
    int sb = 0;                                 // the row's sweep slot of block 0
    if (live) {
        ri = m.rowFwd[I];
        sb = m.swBaseF[I];
    }
    // the flags of one instance, loaded with the row records (the first lane's record start,
    // never negative, joins the test so that those loads are issued before the branch)
    if (!sweep_row_live(m, 0, sc, skip, true) | (__builtin_amdgcn_readfirstlane(ri.x) < 0)) return;
    if (live) {
        if (!sweep_row_live(m, ri.w, sc, skip, false)) {
            if constexpr (!DOT) return;
            live = false;
        }
    }
    double v[3] = {0, 0, 0};
    if (live) {
        const int kb = ri.x, ke = ri.y;
        // the diagonal block and the right-hand side, loaded with the blocks
        T d[6], rhs[3];
        load_diag<T>(U, I, d);
        rhs[0] = r[3 * I]; rhs[1] = r[3 * I + 1]; rhs[2] = r[3 * I + 2];
        T acc[G][3];
#pragma unroll
        for (int g = 0; g < G; g++) acc[g][0] = acc[g][1] = acc[g][2] = 0;
        for (int kk = kb + lane; kk < ke; kk += 16) {
#pragma unroll
            for (int g = 0; g < G; g++) {
                const int k = kk + g * L;
                if (k < ke) {
                    BlockX<T> b;
                    load_factor_block<T>(Ul, m.swNF, sb + (int64_t)(k - kb) * kChunkRows, y + 3 * m.tRow[k], b);
                    acc_add(acc[g][0], row3(b.e[0], b.e[3], b.e[6], b.x[0], b.x[1], b.x[2]));
                    acc_add(acc[g][1], row3(b.e[1], b.e[4], b.e[7], b.x[0], b.x[1], b.x[2]));
                    acc_add(acc[g][2], row3(b.e[2], b.e[5], b.e[8], b.x[0], b.x[1], b.x[2]));
                }
            }
        }
        T a0, a1, a2;
        combine16<L>(acc, a0, a1, a2);
        if (lane == 0) {
            T z0 = (rhs[0] - a0) / d[0];
            T z1 = (rhs[1] - a1 - d[1] * z0) / d[3];
            T z2 = (rhs[2] - a2 - d[2] * z0 - d[4] * z1) / d[5];
            if (lastLevel) diag_bwd(d, z0, z1, z2);   // rows of the last level have no blocks above them
            y[3 * I] = z0; y[3 * I + 1] = z1; y[3 * I + 2] = z2;
            v[0] = (double)rhs[0] * (double)z0; v[1] = (double)rhs[1] * (double)z1; v[2] = (double)rhs[2] * (double)z2;
        }
    }
    if constexpr (DOT) sweep_dot<DOT>(m, c, p0, count, level, lr, lane == 0, v, partials);
}

// backward: U z = y  ->  z_I = U_II^-1 (y_I - sum_{J>I} U_IJ z_J); with DOT r is the residual
// (the r.z terms of the final z)
template <typename T, int L, int DOT>
__global__ __launch_bounds__(DOT ? kChunkRows * L : kIcThreads) void k_ic0_bwd(
    DevMesh m, int p0, int count, const T *__restrict__ U, T *__restrict__ z, const InstScalars *__restrict__ sc,
    const uint8_t *__restrict__ skip, const T *__restrict__ r, double *__restrict__ partials, int level) {
    constexpr int kRows = DOT ? kChunkRows : kIcThreads / L, G = 16 / L;
    const int lane = threadIdx.x % L, lr = threadIdx.x / L;
    int I, c = 0;
    bool live;
    if constexpr (DOT) {
        if constexpr (DOT == 1) {
            const int nwg = sweep_chunks(p0, count), wg = xcd_wg(nwg);
            if (wg >= nwg) return;              // grid padding: the whole workgroup
            c = p0 / kChunkRows + wg;
            I = c * kChunkRows + lr;
            live = I >= p0 && I < p0 + count;
        } else {                                // several instances: p0 and count name the
            const int wg = xcd_wg(count);       // level's range of the chunk list
            if (wg >= count) return;
            c = m.levelChunks[p0 + wg];
            I = m.chunkRow0[c] + lr;
            live = I < m.chunkRow1[c] && m.rowLevel[I] == level;
        }
    } else {
        const int pl = xcd_wg((count + kRows - 1) / kRows) * kRows + lr;
        if (pl >= count) return;
        I = m.levelsContiguous ? p0 + pl : m.levelRows[p0 + pl];
        live = true;
    }
    int4 ri = make_int4(0, 0, 0, 0);
    int sb = 0;                                 // the row's sweep slot of block 0
    if (live) {
        ri = m.rowBwd[I];
        sb = m.swBaseB[I];
    }
    // the flags of one instance, loaded with the row records (the first lane's record start,
    // never negative, joins the test so that those loads are issued before the branch)
    if (!sweep_row_live(m, 0, sc, skip, true) | (__builtin_amdgcn_readfirstlane(ri.x) < 0)) return;
    if (live) {
        if (!sweep_row_live(m, ri.w, sc, skip, false)) {
            if constexpr (!DOT) return;
            live = false;
        }
    }
    double v[3] = {0, 0, 0};
    if (live) {
        const int b0 = ri.x, nOff = ri.y - b0 - 1;
        // the diagonal block and the forward result, loaded with the blocks
        T d[6], yv[3], rv[3] = {0, 0, 0};
        load_diag<T>(U, I, d);
        yv[0] = z[3 * I]; yv[1] = z[3 * I + 1]; yv[2] = z[3 * I + 2];
        if constexpr (DOT) { rv[0] = r[3 * I]; rv[1] = r[3 * I + 1]; rv[2] = r[3 * I + 2]; }
        const T *Uo = U + 8 * (int64_t)m.nRows;   // the part-major upper blocks
        T acc[G][3];
#pragma unroll
        for (int g = 0; g < G; g++) acc[g][0] = acc[g][1] = acc[g][2] = 0;
        for (int kk = lane; kk < nOff; kk += 16) {
#pragma unroll
            for (int g = 0; g < G; g++) {
                const int k = kk + g * L;
                if (k < nOff) {
                    BlockX<T> b;
                    load_factor_block<T>(Uo, m.swNB, sb + (int64_t)k * kChunkRows, z + 3 * m.colIdx[b0 + 1 + k], b);
                    acc_add(acc[g][0], row3(b.e[0], b.e[1], b.e[2], b.x[0], b.x[1], b.x[2]));
                    acc_add(acc[g][1], row3(b.e[3], b.e[4], b.e[5], b.x[0], b.x[1], b.x[2]));
                    acc_add(acc[g][2], row3(b.e[6], b.e[7], b.e[8], b.x[0], b.x[1], b.x[2]));
                }
            }
        }
        T a0, a1, a2;
        combine16<L>(acc, a0, a1, a2);
        if (lane == 0) {
            const T x2 = (yv[2] - a2) / d[5];
            const T x1 = (yv[1] - a1 - d[4] * x2) / d[3];
            const T x0 = (yv[0] - a0 - d[1] * x1 - d[2] * x2) / d[0];
            z[3 * I] = x0; z[3 * I + 1] = x1; z[3 * I + 2] = x2;
            v[0] = (double)rv[0] * (double)x0; v[1] = (double)rv[1] * (double)x1; v[2] = (double)rv[2] * (double)x2;
        }
    }
    if constexpr (DOT) sweep_dot<DOT>(m, c, p0, count, level, lr, lane == 0, v, partials);
}

// Sweeps over the sweep slots (FP32 or FP64 copies): one chunk of 64 rows per workgroup, 16 / K waves.  The
// thread of row r in wave w keeps the virtual lanes v = w + (16 / K) i (i < K); virtual lane v
// takes the row's blocks v, v + 16, ..., each wave load one slot position of all 64 rows
// (1 KB of consecutive float4 or double2 per part).  The sums combine as the 16-lane tree combines
// them (pairs at distance 8, 4, 2, 1: inside the thread first, then across the waves through
// LDS), so every K gives the bits of the lane-per-row kernels above.  Rows of the chunk
// outside the level load their slots and discard the sums.  The x the blocks multiply come
// from the chunk's halo, staged in LDS while the blocks load, so no block waits on an x
// gather.
template <typename T, bool FWD, int K, int DOT>
__global__ __launch_bounds__(1024 / K) void k_ic0_slots(
    DevMesh m, int p0, int count, int level, const int *__restrict__ chunks, const T *__restrict__ U,
    const T *__restrict__ Ul, const T *__restrict__ r, T *__restrict__ y,
    const InstScalars *__restrict__ sc, const uint8_t *__restrict__ skip, int lastLevel, int nwg,
    double *__restrict__ partials) {
    constexpr int W = 16 / K;
    const int wg = xcd_wg(nwg);
    if (wg >= nwg) return;                      // grid padding: the whole workgroup
    // one instance: 64-row chunks from row 0, and with the level a row range its chunks are
    // consecutive (no list load); the chunk's record is loaded with the flags
    const bool one = m.nInst == 1;
    const int c = one && m.levelsContiguous ? p0 / kChunkRows + wg : chunks[wg];
    const int rr = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int4 ch = m.swChunk[c];
    // the chunk's halo rows (fixed stride per chunk: no record needed), loaded with the record
    constexpr int NT = 1024 / K, HR = 2;
    const int2 hcn = m.swHaloCnt[c];
    const int H = FWD ? m.swHF : m.swHB;
    const int *__restrict__ hrows = (FWD ? m.swHaloF : m.swHaloB) + (int64_t)c * H;
    int hr[HR];
#pragma unroll
    for (int i = 0; i < HR; i++) {
        const int h = (int)threadIdx.x + i * NT;
        hr[i] = hrows[h < H ? h : 0];
    }
    const int row0 = one ? c * kChunkRows : m.chunkRow0[c], I = row0 + rr;
    const int row1 = one ? min(row0 + kChunkRows, m.nRows) : m.chunkRow1[c];
    const int off = FWD ? ch.x : ch.z, width = FWD ? ch.y : ch.w;
    // (offset and width, never negative, join the test so that their load is issued with the
    // flags' instead of after the branch)
    if (!sweep_row_live(m, one ? 0 : m.chunkInst[c], sc, skip, one) | ((off | width) < 0)) return;
    const bool live = I < row1 && (m.levelsContiguous ? I >= p0 && I < p0 + count : m.rowLevel[I] == level);
    const int64_t nb = FWD ? m.swNF : m.swNB;
    const T *P = FWD ? Ul : U + 8 * (int64_t)m.nRows;
    const int *__restrict__ colA = FWD ? m.swColF : m.swColB;
    // the diagonal block and the right-hand side (forward: r; backward: the forward result),
    // loaded with the blocks
    T d[6] = {1, 0, 0, 1, 0, 1}, rhs[3] = {0, 0, 0}, rv[3] = {0, 0, 0};
    if (w == 0 && live) {
        load_diag<T>(U, I, d);
        const T *src = FWD ? r : y;
        rhs[0] = src[3 * I]; rhs[1] = src[3 * I + 1]; rhs[2] = src[3 * I + 2];
        if constexpr (DOT && !FWD) { rv[0] = r[3 * I]; rv[1] = r[3 * I + 1]; rv[2] = r[3 * I + 2]; }
    }
    T acc[K][3];
#pragma unroll
    for (int i = 0; i < K; i++) acc[i][0] = acc[i][1] = acc[i][2] = 0;
    T e[K][9];
    int hs[K];
    bool use[K];
    // every load first: positions past the width repeat the last one (in the chunk's slots),
    // padding columns name halo slot 0
    auto load = [&](int kb) {
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int k = kb + w + W * i;
            const int64_t pos = off + (int64_t)min(k, width - 1) * kChunkRows + rr;
            const int col = colA[pos];
            use[i] = live && k < width && col >= 0;
            load_pm(P, nb, pos, e[i]);
            hs[i] = max(col, 0);
        }
    };
    if (width > 0) load(0);
    // the halo's x into LDS (y of earlier levels, final; a row of this launch's level is
    // named only by rows outside the level, whose sums are dropped)
    extern __shared__ __align__(16) unsigned char swLds[];
    T *hx = reinterpret_cast<T *>(swLds);
    const int hc = FWD ? hcn.x : hcn.y;   // the chunk's own halo rows (the rest is padding)
#pragma unroll
    for (int i = 0; i < HR; i++) {
        const int h = (int)threadIdx.x + i * NT;
        if (h < hc) { hx[3 * h] = y[3 * hr[i]]; hx[3 * h + 1] = y[3 * hr[i] + 1]; hx[3 * h + 2] = y[3 * hr[i] + 2]; }
    }
    for (int h = (int)threadIdx.x + HR * NT; h < hc; h += NT) {
        const int hrw = hrows[h];
        hx[3 * h] = y[3 * hrw]; hx[3 * h + 1] = y[3 * hrw + 1]; hx[3 * h + 2] = y[3 * hrw + 2];
    }
    __syncthreads();
    for (int kb = 0; kb < width; kb += 16) {
        if (kb > 0) load(kb);
        T xv[K][3];
#pragma unroll
        for (int i = 0; i < K; i++) { xv[i][0] = hx[3 * hs[i]]; xv[i][1] = hx[3 * hs[i] + 1]; xv[i][2] = hx[3 * hs[i] + 2]; }
#pragma unroll
        for (int i = 0; i < K; i++) {
            // forward: U_KI^T y_K (columns of the stored block); backward: U_IJ z_J (rows)
            const T t0 = FWD ? row3(e[i][0], e[i][3], e[i][6], xv[i][0], xv[i][1], xv[i][2])
                                 : row3(e[i][0], e[i][1], e[i][2], xv[i][0], xv[i][1], xv[i][2]);
            const T t1 = FWD ? row3(e[i][1], e[i][4], e[i][7], xv[i][0], xv[i][1], xv[i][2])
                                 : row3(e[i][3], e[i][4], e[i][5], xv[i][0], xv[i][1], xv[i][2]);
            const T t2 = FWD ? row3(e[i][2], e[i][5], e[i][8], xv[i][0], xv[i][1], xv[i][2])
                                 : row3(e[i][6], e[i][7], e[i][8], xv[i][0], xv[i][1], xv[i][2]);
            T a0 = acc[i][0], a1 = acc[i][1], a2 = acc[i][2];
            acc_add(a0, t0); acc_add(a1, t1); acc_add(a2, t2);
            if (use[i]) { acc[i][0] = a0; acc[i][1] = a1; acc[i][2] = a2; }
        }
    }
    // the tree inside the thread: virtual distance dv is i-distance dv / W
#pragma unroll
    for (int dv = 8; dv >= W; dv >>= 1)
#pragma unroll
        for (int i = 0; i < dv / W; i++) {
            acc_add(acc[i][0], acc[i + dv / W][0]);
            acc_add(acc[i][1], acc[i + dv / W][1]);
            acc_add(acc[i][2], acc[i + dv / W][2]);
        }
    T a0 = acc[0][0], a1 = acc[0][1], a2 = acc[0][2];
    if constexpr (W > 1) {
        __shared__ T red[W][kChunkRows][3];
        red[w][rr][0] = a0; red[w][rr][1] = a1; red[w][rr][2] = a2;
        __syncthreads();
        if (w == 0) {
            T t[W][3];
#pragma unroll
            for (int j = 0; j < W; j++) { t[j][0] = red[j][rr][0]; t[j][1] = red[j][rr][1]; t[j][2] = red[j][rr][2]; }
#pragma unroll
            for (int n = W; n > 1; n >>= 1)
#pragma unroll
                for (int j = 0; j < n / 2; j++) {
                    acc_add(t[j][0], t[j + n / 2][0]);
                    acc_add(t[j][1], t[j + n / 2][1]);
                    acc_add(t[j][2], t[j + n / 2][2]);
                }
            a0 = t[0][0]; a1 = t[0][1]; a2 = t[0][2];
        }
    }
    double v[3] = {0, 0, 0};
    if (w == 0 && live) {
        if constexpr (FWD) {
            T z0 = (rhs[0] - a0) / d[0];
            T z1 = (rhs[1] - a1 - d[1] * z0) / d[3];
            T z2 = (rhs[2] - a2 - d[2] * z0 - d[4] * z1) / d[5];
            if (lastLevel) diag_bwd(d, z0, z1, z2);   // rows of the last level have no blocks above them
            y[3 * I] = z0; y[3 * I + 1] = z1; y[3 * I + 2] = z2;
            v[0] = (double)rhs[0] * (double)z0; v[1] = (double)rhs[1] * (double)z1; v[2] = (double)rhs[2] * (double)z2;
        } else {
            const T x2 = (rhs[2] - a2) / d[5];
            const T x1 = (rhs[1] - a1 - d[4] * x2) / d[3];
            const T x0 = (rhs[0] - a0 - d[1] * x1 - d[2] * x2) / d[0];
            y[3 * I] = x0; y[3 * I + 1] = x1; y[3 * I + 2] = x2;
            v[0] = (double)rv[0] * (double)x0; v[1] = (double)rv[1] * (double)x1; v[2] = (double)rv[2] * (double)x2;
        }
    }
    if constexpr (DOT) sweep_dot<DOT>(m, c, p0, count, level, rr, w == 0, v, partials);
}

// One sweep launch of level l in the form `code` names: code & 31 is the lanes per row of
// the lane-per-row kernels (4, 8 or 16), code >> 5 (1-4) selects the slot kernel with K = 2,
// 4, 8 or 16 slots per thread.  With partials set, the launch
// also forms the r.z chunk partials of the rows whose z it finishes (one instance,
// consecutive levels).  The optional events are recorded from the dispatch's own begin/end
// timestamps (lane autotuner).
template <typename T>
static void launch_sweep(bool fwd, int code, const DevMesh &m, hipStream_t s, int level, int p0, int cnt, const T *U,
                         const T *Ul, const T *r, T *z, const InstScalars *sc, const uint8_t *skip, int last,
                         double *partials, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const bool dot = partials != nullptr;
    if (dot && !(m.nInst == 1 ? m.levelsContiguous : m.sweepDotLists))
        throw HipError("sweep dot needs one instance with consecutive levels, or short level chunk lists");
    const int L = code & 31, sel = code >> 5;
    if (sel) {
        const int c0 = m.levelChunkPtr[level], nwg = m.levelChunkPtr[level + 1] - c0;
        if (nwg <= 0) return;
        const int K = 1 << sel;
        const size_t lds = 3 * sizeof(T) * (size_t)(fwd ? m.swHF : m.swHB);   // the chunk's halo
        auto go = [&](auto kernel) {
            const dim3 grid(xcd_grid(nwg)), block(1024 / K);
            static bool raised = false;   // a halo beyond 48 KB: the kernel's dynamic LDS limit raised once
            if (lds > 49152 && !raised) {
                OD_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds_max() - 16384));
                raised = true;
            }
            if (e0) hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e0, e1, 0, m, p0, cnt, level,
                                          m.levelChunks + c0, U, Ul, r, z, sc, skip, last, nwg, partials);
            else hipLaunchKernelGGL(kernel, grid, block, lds, s, m, p0, cnt, level, m.levelChunks + c0, U, Ul, r, z,
                                    sc, skip, last, nwg, partials);
        };
        auto pick = [&](auto fwdTag, auto dotTag) {
            constexpr bool F = decltype(fwdTag)::value;
            constexpr int D = decltype(dotTag)::value;
            if (K == 2) go(k_ic0_slots<T, F, 2, D>);
            else if (K == 4) go(k_ic0_slots<T, F, 4, D>);
            else if (K == 8) go(k_ic0_slots<T, F, 8, D>);
            else go(k_ic0_slots<T, F, 16, D>);
        };
        using One = std::integral_constant<int, 1>;
        using Many = std::integral_constant<int, 2>;
        using None = std::integral_constant<int, 0>;
        if (fwd) {
            if (!dot) pick(std::true_type{}, None{});
            else if (m.nInst == 1) pick(std::true_type{}, One{});
            else pick(std::true_type{}, Many{});
        } else {
            if (!dot) pick(std::false_type{}, None{});
            else if (m.nInst == 1) pick(std::false_type{}, One{});
            else pick(std::false_type{}, Many{});
        }
        return;
    }
    // a dot launch over several instances covers the level's chunks: its range of the chunk list
    const bool chunkList = dot && m.nInst != 1;
    const int q0 = chunkList ? m.levelChunkPtr[level] : p0;
    const int qn = chunkList ? m.levelChunkPtr[level + 1] - q0 : cnt;
    if (chunkList && qn <= 0) return;
    auto go = [&](auto kernel, int lanes, auto... args) {
        const int nwg = chunkList ? qn
                        : dot    ? (p0 + cnt - 1) / kChunkRows - p0 / kChunkRows + 1
                                 : (cnt + kIcThreads / lanes - 1) / (kIcThreads / lanes);
        const dim3 grid(xcd_grid(nwg)), block(dot ? kChunkRows * lanes : kIcThreads);
        if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, m, q0, qn, args...);
        else hipLaunchKernelGGL(kernel, grid, block, 0, s, m, q0, qn, args...);
    };
    auto pick = [&](auto dotTag) {
        constexpr int D = decltype(dotTag)::value;
        if (fwd) {
            if (L == 16) go(k_ic0_fwd<T, 16, D>, 16, U, Ul, r, z, sc, skip, last, partials, level);
            else if (L == 4) go(k_ic0_fwd<T, 4, D>, 4, U, Ul, r, z, sc, skip, last, partials, level);
            else go(k_ic0_fwd<T, 8, D>, 8, U, Ul, r, z, sc, skip, last, partials, level);
        } else {
            if (L == 16) go(k_ic0_bwd<T, 16, D>, 16, U, z, sc, skip, r, partials, level);
            else if (L == 4) go(k_ic0_bwd<T, 4, D>, 4, U, z, sc, skip, r, partials, level);
            else go(k_ic0_bwd<T, 8, D>, 8, U, z, sc, skip, r, partials, level);
        }
    };
    if (!dot) pick(std::integral_constant<int, 0>{});
    else if (m.nInst == 1) pick(std::integral_constant<int, 1>{});
    else pick(std::integral_constant<int, 2>{});
}

// Sweep launches for forward levels [f0, nl) (the last one also solving its backward step
// when fuseLast) and backward levels [0, b1), with each level's lanes per row.
// With dotPartials (fuseLast and b1 = nl - 1: every level's z finishes in one of these
// launches), the launches that finish z also form the r.z chunk partials.
template <typename T>
static void ic0_levels(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, int f0, bool fuseLast, int b1,
                       const T *U, const T *Ul, const T *r, T *z, const InstScalars *sc, const uint8_t *skip,
                       double *dotPartials = nullptr) {
    const int nl = (int)levelPtr.size() - 1;
    for (int l = f0; l < nl; l++) {
        const int cnt = levelPtr[l + 1] - levelPtr[l];
        if (cnt <= 0) continue;
        const int last = (fuseLast && l == nl - 1) ? 1 : 0;
        launch_sweep<T>(true, (sizeof(T) == 8 ? m.levelLanesFwd64 : m.levelLanesFwd)[l], m, s, l, levelPtr[l], cnt, U,
                        Ul, r, z, sc, skip, last, last ? dotPartials : nullptr);
    }
    for (int l = b1 - 1; l >= 0; l--) {
        const int cnt = levelPtr[l + 1] - levelPtr[l];
        if (cnt <= 0) continue;
        launch_sweep<T>(false, (sizeof(T) == 8 ? m.levelLanesBwd64 : m.levelLanesBwd)[l], m, s, l, levelPtr[l], cnt, U,
                        Ul, r, z, sc, skip, 0, dotPartials);
    }
    OD_HIP(hipGetLastError());
}

// Reads a large buffer once (one partial per workgroup): evicts the memory-side cache without
// leaving dirty lines behind, so that a timed launch after it reads its data from HBM, as the
// solver's launches do (one PCG iteration's data exceeds the 256 MB memory-side cache).
__global__ void k_stream_read(const float4 *p, int64_t n, float *out) {
    float s = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = p[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == 12345.f) out[blockIdx.x] = s;   // keeps the loads; the buffer never holds this sum
}

void launch_stream_read(hipStream_t s, float *buf, int64_t nFloats) {
    hipLaunchKernelGGL(k_stream_read, dim3(2048), dim3(256), 0, s, (const float4 *)buf, nFloats / 4, buf);
    OD_HIP(hipGetLastError());
}

// Launch form of each level's sweeps in precision T: times 4, 8 and 16 lanes per row and the
// slot kernel with 2, 4, 8 and 16 slots per thread (median of three launches each, on
// zero-filled factors of this mesh's pattern) and keeps the fastest (launch_sweep's code).
// Each launch finds its data as the solver's sweeps find it: when a PCG iteration's factor
// copies and matrix values exceed the 256 MB memory-side cache, after a read that evicts the
// caches, else warm.  The last level's forward launch is timed with its fused backward step,
// as the solver runs it.
template <typename T>
void tune_ic0_lanes(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, uint8_t *fwd, uint8_t *bwd) {
    const int nl = (int)levelPtr.size() - 1;
    const size_t n = 3 * (size_t)m.nRows;
    const bool cold = (int64_t)sizeof(T) * (ic0_u32_floats(m) + ic0_ul32_floats(m) + m.nValues) > ((int64_t)192 << 20);
    const int64_t flushN = cold ? (int64_t)128 << 20 : 4;    // floats: 512 MB
    std::vector<void *> bufs;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto get = [&](size_t bytes) {
        void *p = nullptr;
        OD_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        bufs.push_back(p);
        OD_HIP(hipMemsetAsync(p, 0, std::max<size_t>(bytes, 16), s));
        return p;
    };
    try {
        T *U = (T *)get(sizeof(T) * (size_t)ic0_u32_floats(m)), *Ul = (T *)get(sizeof(T) * (size_t)ic0_ul32_floats(m));
        T *r = (T *)get(sizeof(T) * n), *z = (T *)get(sizeof(T) * n);
        float *flush = (float *)get(4 * (size_t)flushN);
        // the launches that finish z form the r.z partials when the solver fuses that dot
        double *dots = ic0_dot_fused(m, levelPtr) ? (double *)get(8 * (size_t)std::max(1, m.nChunks)) : nullptr;
        OD_HIP(hipEventCreate(&e0));
        OD_HIP(hipEventCreate(&e1));
        // lane-per-row kernels (4, 8, 16 lanes), then the slot kernel (K = 2, 4, 8, 16): the
        // code keeps the fastest lane count and, when a slot shape beats it, that shape
        const int codes[7] = {4, 8, 16, 1 << 5, 2 << 5, 3 << 5, 4 << 5};
        for (int dir = 0; dir < 2; dir++)
            for (int l = 0; l < nl; l++) {
                const int cnt = levelPtr[l + 1] - levelPtr[l];
                if (cnt <= 0) continue;
                float best = 0.f;
                uint8_t &code = (dir == 0 ? fwd : bwd)[l];
                for (int v = 0; v < 7; v++) {
                    const int cv = v < 3 ? codes[v] : ((code & 31) | codes[v]);
                    float ms[3];
                    for (int rep = 0; rep < 3; rep++) {
                        const bool last = dir == 0 && l == nl - 1, finishes = last || (dir == 1 && l < nl - 1);
                        if (cold) launch_stream_read(s, flush, flushN);
                        launch_sweep<T>(dir == 0, cv, m, s, l, levelPtr[l], cnt, U, Ul, r, z, nullptr, nullptr,
                                        last ? 1 : 0, finishes ? dots : nullptr, e0, e1);
                        OD_HIP(hipGetLastError());
                        OD_HIP(hipEventSynchronize(e1));
                        OD_HIP(hipEventElapsedTime(&ms[rep], e0, e1));
                    }
                    std::sort(ms, ms + 3);
                    if (v == 0 || ms[1] < best) {
                        best = ms[1];
                        code = (uint8_t)cv;
                    }
                }
            }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void *p : bufs) (void)hipFree(p);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (void *p : bufs) (void)hipFree(p);
}

template <typename T>
void launch_ic0_apply(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, const T *U, const T *Ul,
                      const T *r, T *z, const InstScalars *sc, const uint8_t *skip) {
    const int nl = (int)levelPtr.size() - 1;
    ic0_levels<T>(m, s, levelPtr, 0, false, nl, U, Ul, r, z, sc, skip);
}

// x += alpha p ; r -= alpha q ; slot [1] r.r ; then, for the rows of level 0 (no blocks below
// them), the forward solve y = U_II^-T r (and the backward one when there is a single level)
template <typename T>
__global__ __launch_bounds__(kVecThreads) void k_update_xr_ic0(DevMesh m, const T *p, const T *q, T *x, T *r,
                                                               const T *__restrict__ U, T *z, int singleLevel,
                                                               int level0End, double *partials,
                                                               const InstScalars *sc) {
    CHUNK_PROLOGUE
    __shared__ T rs[kVecThreads];
    const bool act = !sc[inst].done;
    // rows of level 0 (rows below level0End when the levels are consecutive ranges): their
    // diagonal blocks are loaded with the vectors
    const int row = row0 + (int)threadIdx.x;
    const bool lev0 = act && threadIdx.x < kChunkRows && row < row1 &&
                      (level0End >= 0 ? row < level0End : m.rowLevel[row] == 0);
    T d[6] = {1, 0, 0, 1, 0, 1};
    if (lev0) load_diag<T>(U, row, d);
    double v[1] = {0};
    T ri = 0;
    if (in && act) {
        const T a = (T)sc[inst].alpha;
        x[i] = x[i] + a * p[i];
        ri = r[i] - a * q[i];
        r[i] = ri;
        v[0] = (double)ri * (double)ri;
    }
    rs[threadIdx.x] = ri;
    block_sum<1>(v);                       // its barrier also publishes rs
    if (threadIdx.x == 0) partials[m.nChunks + chunk] = v[0];
    if (!lev0) return;
    const int b = 3 * (int)threadIdx.x;
    T y0 = rs[b] / d[0];
    T y1 = (rs[b + 1] - d[1] * y0) / d[3];
    T y2 = (rs[b + 2] - d[2] * y0 - d[4] * y1) / d[5];
    if (singleLevel) diag_bwd(d, y0, y1, y2);
    z[3 * row] = y0; z[3 * row + 1] = y1; z[3 * row + 2] = y2;
}

template <typename T>
void pcg_update_xr_ic0(const DevMesh &m, hipStream_t s, const std::vector<int> &levelPtr, const T *p, const T *q, T *x,
                       T *r, const T *U, const T *Ul, T *z, PcgWork &w, const uint8_t *skip) {
    const int nl = (int)levelPtr.size() - 1;
    hipLaunchKernelGGL(k_update_xr_ic0<T>, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, p, q, x, r, U, z,
                       nl == 1 ? 1 : 0, m.levelsContiguous && nl >= 1 ? levelPtr[1] : -1, w.partials, w.sc);
    OD_HIP(hipGetLastError());
    if (nl > 1)
        ic0_levels<T>(m, s, levelPtr, 1, true, nl - 1, U, Ul, r, z, w.sc, skip,
                      ic0_dot_fused(m, levelPtr) ? w.partials : nullptr);
}

bool ic0_dot_fused(const DevMesh &m, const std::vector<int> &levelPtr) {
    return (m.nInst == 1 ? m.levelsContiguous : m.sweepDotLists) && levelPtr.size() >= 3;
}

void pcg_finalize_iter(const DevMesh &m, hipStream_t s, PcgWork &w) {
    launch_finalize(m, s, PH_ITER, w, 0.0, 0, nullptr);
}

// =====================================================================================
// mixed precision and integrator updates
// =====================================================================================
__global__ void k_cast(const double *a, float *b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
void launch_cast_f64_f32(hipStream_t s, const double *a, float *b, int64_t n) {
    hipLaunchKernelGGL(k_cast, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a, b, n);
    OD_HIP(hipGetLastError());
}

// r = b - Ax (FP64), r32 = (float) r ; slots [1] r.r, [2] b.b ; then the outer done flags
__global__ __launch_bounds__(kVecThreads) void k_outer_res(DevMesh m, const double *b, const double *Ax, double *r,
                                                           float *r32, double *partials) {
    CHUNK_PROLOGUE
    double v[2] = {0, 0};
    if (in) {
        double ri = b[i] - Ax[i];
        r[i] = ri;
        r32[i] = (float)ri;
        v[0] = ri * ri;
        v[1] = b[i] * b[i];
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) { partials[m.nChunks + chunk] = v[0]; partials[2 * m.nChunks + chunk] = v[1]; }
}

void launch_residual_f64(const DevMesh &m, hipStream_t s, const double *b, const double *Ax, double *r, float *r32,
                         PcgWork &w, double tol, const uint8_t *preDone, int first) {
    hipLaunchKernelGGL(k_outer_res, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, b, Ax, r, r32, w.partials);
    OD_HIP(hipGetLastError());
    launch_finalize(m, s, first ? PH_OUTER_FIRST : PH_OUTER, w, tol, 0, preDone);
}

__global__ __launch_bounds__(kVecThreads) void k_add_corr(DevMesh m, const float *dx, double *x, const InstScalars *outer) {
    CHUNK_PROLOGUE
    if (!in || outer[inst].done) return;
    x[i] += (double)dx[i];
}
void launch_add_correction(const DevMesh &m, hipStream_t s, const float *dx, double *x, const InstScalars *outer) {
    hipLaunchKernelGGL(k_add_corr, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, dx, x, outer);
    OD_HIP(hipGetLastError());
}

// Newmark corrector (linearNewmark.cpp:68-69, 76-77), predictor recomputed from the old state
__global__ __launch_bounds__(kVecThreads) void k_newmark_correct(DevMesh m, double *d, double *v, const double *aOld,
                                                                 const double *aNew, double h, double beta,
                                                                 double gamma, const uint8_t *instActive) {
    CHUNK_PROLOGUE
    if (!in || (instActive && !instActive[inst])) return;
    const double bh2 = beta * h * h, gh = gamma * h, mbh2 = h * h * 0.5 - bh2, mgh = h - gh;
    double ao = aOld[i], an = aNew[i];
    double dp = d[i] + h * v[i] + mbh2 * ao;
    double vp = v[i] + mgh * ao;
    v[i] = vp + gh * an;
    d[i] = dp + bh2 * an;
}
void launch_newmark_correct(const DevMesh &m, hipStream_t s, double *d, double *v, const double *aOld,
                            const double *aNew, double h, double beta, double gamma, const uint8_t *instActive) {
    hipLaunchKernelGGL(k_newmark_correct, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, d, v, aOld, aNew, h, beta, gamma,
                       instActive);
    OD_HIP(hipGetLastError());
}

// k_newmark_correct and the copy a = aNew in one pass, for the instances `commit` marks
__global__ __launch_bounds__(kVecThreads) void k_newmark_commit(DevMesh m, double *d, double *v, double *a,
                                                                const double *aNew, double h, double beta,
                                                                double gamma, const uint8_t *commit) {
    CHUNK_PROLOGUE
    if (!in || !commit[inst]) return;
    const double bh2 = beta * h * h, gh = gamma * h, mbh2 = h * h * 0.5 - bh2, mgh = h - gh;
    double ao = a[i], an = aNew[i];
    double dp = d[i] + h * v[i] + mbh2 * ao;
    double vp = v[i] + mgh * ao;
    v[i] = vp + gh * an;
    d[i] = dp + bh2 * an;
    a[i] = an;
}
void launch_newmark_commit(const DevMesh &m, hipStream_t s, double *d, double *v, double *a, const double *aNew,
                           double h, double beta, double gamma, const uint8_t *commit) {
    hipLaunchKernelGGL(k_newmark_commit, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, d, v, a, aNew, h, beta, gamma,
                       commit);
    OD_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kVecThreads) void k_axpy_rows(DevMesh m, double *y, double a, const double *x,
                                                           const uint8_t *instActive) {
    CHUNK_PROLOGUE
    if (!in || (instActive && !instActive[inst])) return;
    y[i] += a * x[i];
}
void launch_axpy_rows(const DevMesh &m, hipStream_t s, double *y, double a, const double *x, const uint8_t *instActive) {
    hipLaunchKernelGGL(k_axpy_rows, dim3(m.nChunks), dim3(kVecThreads), 0, s, m, y, a, x, instActive);
    OD_HIP(hipGetLastError());
}
void launch_be_correct(const DevMesh &m, hipStream_t s, double *v, const double *dv, const uint8_t *instActive) {
    launch_axpy_rows(m, s, v, 1.0, dv, instActive);
}

__global__ void k_done_from_active(int nInst, uint8_t *preDone, const uint8_t *active) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nInst) preDone[i] = active ? (active[i] ? 0 : 1) : 0;
}
void launch_set_done_from_active(const DevMesh &m, hipStream_t s, uint8_t *preDone, const uint8_t *instActive) {
    hipLaunchKernelGGL(k_done_from_active, dim3((m.nInst + 255) / 256), dim3(256), 0, s, m.nInst, preDone, instActive);
    OD_HIP(hipGetLastError());
}

// explicit instantiations
#define INST_T(T)                                                                                                     \
    template void pcg_init_residual<T>(const DevMesh &, hipStream_t, const T *, const T *, T *, PcgWork &);           \
    template void pcg_jacobi<T>(const DevMesh &, hipStream_t, const T *, const T *, T *, PcgWork &);                  \
    template void pcg_dot_rz_init<T>(const DevMesh &, hipStream_t, const T *, const T *, T *, PcgWork &, double, int, \
                                     const uint8_t *);                                                                \
    template void pcg_alpha<T>(const DevMesh &, hipStream_t, const T *, const T *, const T *, T *, T *, PcgWork &,    \
                               hipEvent_t, hipEvent_t);                                                           \
    template void pcg_update_xr<T>(const DevMesh &, hipStream_t, const T *, const T *, T *, T *, const T *, T *,      \
                                   PcgWork &);                                                                        \
    template void pcg_dot_rz<T>(const DevMesh &, hipStream_t, const T *, const T *, PcgWork &);                       \
    template void pcg_update_p<T>(const DevMesh &, hipStream_t, const T *, T *, PcgWork &);                           \
    template void pcg_update_xr_ic0<T>(const DevMesh &, hipStream_t, const std::vector<int> &, const T *, const T *,   \
                                       T *, T *, const T *, const T *, T *, PcgWork &, const uint8_t *);              \
    template void launch_jacobi_diag<T>(const DevMesh &, hipStream_t, const double *, T *);                           \
    template void launch_ic0_apply<T>(const DevMesh &, hipStream_t, const std::vector<int> &, const T *, const T *,   \
                                      const T *, T *, const InstScalars *, const uint8_t *);
INST_T(double)
INST_T(float)
template void launch_ic0_factor<double>(const DevMesh &, hipStream_t, const std::vector<int> &, const double *, double *,
                                        double *, double *, double, const uint8_t *);
template void launch_ic0_factor<float>(const DevMesh &, hipStream_t, const std::vector<int> &, const double *, double *,
                                       float *, float *, double, const uint8_t *);
template void tune_ic0_lanes<double>(const DevMesh &, hipStream_t, const std::vector<int> &, uint8_t *, uint8_t *);
template void tune_ic0_lanes<float>(const DevMesh &, hipStream_t, const std::vector<int> &, uint8_t *, uint8_t *);

}  // namespace od
