// Device-side per-tet numerics (FP64, one thread per tet).
//
// The branch structure of each routine follows the reference routine it names,
// because the element values must agree with the reference's CPU path:
//   polar3     core/numerMethod.h:570-1016   (rotation via the dominant eigenvector
//                                             of the 4x4 quaternion matrix)
//   qr3        core/numerMethod.h:431-507   (Givens, proper rotation)
//   eigsym3    core/numerMethod.h:327-429   (closed-form symmetric eigenpairs)
//   tet_dn     element/tetelement.cpp:18-48 (6V * grad N with orientation fix)
//   neo_*      material/invertibleHyperelastic.cpp:116-265,
//              material/invertibleNeoHookean.cpp:11-20, element/tetelement.cpp:282-499
// Matrices are row-major double[9].
#pragma once
#include <hip/hip_runtime.h>

namespace odk {

struct v3 { double x, y, z; };
__device__ __forceinline__ v3 mk(double a, double b, double c) { v3 r; r.x = a; r.y = b; r.z = c; return r; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator*(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ v3 vdiv(v3 a, double f) { double i = 1.0 / f; return mk(a.x * i, a.y * i, a.z * i); }
__device__ __forceinline__ double comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

__device__ __forceinline__ double det3(const double *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
__device__ __forceinline__ void inv3(const double *t, double *r) {
    double invDet = 1.0 / det3(t);
    r[0] = (t[4] * t[8] - t[5] * t[7]) * invDet;
    r[1] = (t[2] * t[7] - t[1] * t[8]) * invDet;
    r[2] = (t[1] * t[5] - t[2] * t[4]) * invDet;
    r[3] = (t[5] * t[6] - t[3] * t[8]) * invDet;
    r[4] = (t[0] * t[8] - t[2] * t[6]) * invDet;
    r[5] = (t[2] * t[3] - t[0] * t[5]) * invDet;
    r[6] = (t[3] * t[7] - t[4] * t[6]) * invDet;
    r[7] = (t[1] * t[6] - t[0] * t[7]) * invDet;
    r[8] = (t[0] * t[4] - t[1] * t[3]) * invDet;
}
__device__ __forceinline__ void mul3(const double *a, const double *b, double *c) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
}
__device__ __forceinline__ v3 mulv(const double *m, v3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

__device__ __forceinline__ double tet_volume(v3 a, v3 b, v3 c, v3 d) {
    return fabs(dot(b - a, cross(c - a, d - a))) / 6.0;
}

__device__ __forceinline__ void tet_dn(v3 a, v3 b, v3 c, v3 d, v3 *dn) {
    v3 da = a - d, db = b - d, dc = c - d, ca = a - c, cb = b - c;
    if (dot(da, cross(db, dc)) > 0.0) {
        dn[0] = cross(db, dc); dn[1] = cross(dc, da); dn[2] = cross(da, db); dn[3] = cross(cb, ca);
    } else {
        dn[0] = cross(dc, db); dn[1] = cross(da, dc); dn[2] = cross(db, da); dn[3] = cross(ca, cb);
    }
}

// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void qr3(const double *mat, double *ortho) {
    double Q[9], R[9];
#pragma unroll
    for (int i = 0; i < 9; i++) { R[i] = mat[i]; Q[i] = 0.0; }
    double c = 0, s = 0, r = 0;
    auto givens = [](double a, double b, double &c, double &s, double &r) {
        if (b == 0.0) { c = a < 0.0 ? -1.0 : 1.0; s = 0.0; r = fabs(a); }
        else if (a == 0.0) { c = 0.0; s = b < 0.0 ? 1.0 : -1.0; r = fabs(b); }
        else if (fabs(b) > fabs(a)) {
            double tau = -a / b, u = sqrt(1.0 + tau * tau);
            if (b < 0.0) u = -u;
            s = -1.0 / u; c = s * tau; r = b * u;
        } else {
            double tau = -b / a, u = sqrt(1.0 + tau * tau);
            if (a < 0.0) u = -u;
            c = 1.0 / u; s = c * tau; r = a * u;
        }
    };
    givens(R[3], R[6], c, s, r);
    Q[4] = Q[8] = c; Q[5] = s; Q[7] = -s;
    R[3] = r; R[6] = 0.0;
    for (int i = 0; i < 2; i++) { double p = c * R[4 + i] - s * R[7 + i], q = s * R[4 + i] + c * R[7 + i]; R[4 + i] = p; R[7 + i] = q; }
    givens(R[0], R[3], c, s, r);
    R[0] = r; R[3] = 0.0;
    for (int i = 0; i < 2; i++) { double p = c * R[1 + i] - s * R[4 + i], q = s * R[1 + i] + c * R[4 + i]; R[1 + i] = p; R[4 + i] = q; }
    Q[0] = c; Q[1] = s;
    Q[3] = Q[4] * -s; Q[4] *= c;
    Q[6] = Q[7] * -s; Q[7] *= c;
    givens(R[4], R[7], c, s, r);
    for (int i = 0; i < 3; i++) { double p = c * Q[i * 3 + 1] - s * Q[i * 3 + 2], q = s * Q[i * 3 + 1] + c * Q[i * 3 + 2]; Q[i * 3 + 1] = p; Q[i * 3 + 2] = q; }
#pragma unroll
    for (int i = 0; i < 9; i++) ortho[i] = Q[i];
}

// Householder pair used by the small-omega branch of polar3 (numerMethod.h:509-568)
__device__ __forceinline__ void orthonormalize8(double *mat) {
    const double tol = 2.2250738585072014e-308;
    double beta0 = 0, beta1 = 0, v0[3], v1[2];
    v0[0] = mat[1]; v0[1] = mat[2]; v0[2] = mat[3];
    double n2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    if (n2 <= tol) { v0[0] = v0[1] = v0[2] = 0; }
    else {
        double c0 = mat[0], sigma = sqrt(c0 * c0 + n2);
        if (c0 >= 0.0) sigma = -sigma;
        double f = c0 - sigma;
        v0[0] /= f; v0[1] /= f; v0[2] /= f;
        beta0 = (sigma - c0) / sigma;
    }
    double product = mat[4] + v0[0] * mat[5] + v0[1] * mat[6] + v0[2] * mat[7];
    double c0 = mat[5] - product * beta0 * v0[0];
    v1[0] = mat[6] - product * beta0 * v0[1];
    v1[1] = mat[7] - product * beta0 * v0[2];
    n2 = v1[0] * v1[0] + v1[1] * v1[1];
    if (n2 <= tol) { v1[0] = v1[1] = 0; }
    else {
        double sigma = sqrt(c0 * c0 + n2);
        if (c0 >= 0.0) sigma = -sigma;
        double f = c0 - sigma;
        v1[0] /= f; v1[1] /= f;
        beta1 = (sigma - c0) / sigma;
    }
    mat[0] = 1.0 - beta0; mat[1] = -beta0 * v0[0]; mat[2] = -beta0 * v0[1]; mat[3] = -beta0 * v0[2];
    double s0 = 1.0 - beta0 * v0[0] * v0[0], s1 = -beta0 * v0[0] * v0[1], s2 = -beta0 * v0[0] * v0[2];
    double s4 = 1.0 - beta0 * v0[1] * v0[1], s5 = -beta0 * v0[1] * v0[2], s8 = 1.0 - beta0 * v0[2] * v0[2];
    double l0 = 1.0 - beta1, l1 = -beta1 * v1[0], l2 = -beta1 * v1[1];
    mat[4] = mat[1] * l0 + mat[2] * l1 + mat[3] * l2;
    mat[5] = s0 * l0 + s1 * l1 + s2 * l2;
    mat[6] = s1 * l0 + s4 * l1 + s5 * l2;
    mat[7] = s2 * l0 + s5 * l1 + s8 * l2;
}

// Rotation factor R of F = R S (R returned row-major; the reference's `ortho`).
// Symmetric swap of positions K and J of a full 4x4 matrix (rows, then columns) and of the
// index record ix, when c holds; K and J are compile-time, so every entry stays in registers.
template <int K, int J>
__device__ __forceinline__ void sym_swap4(double (&P)[16], int (&ix)[4], bool c) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const double a = P[K * 4 + t], b = P[J * 4 + t];
        P[K * 4 + t] = c ? b : a; P[J * 4 + t] = c ? a : b;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const double a = P[t * 4 + K], b = P[t * 4 + J];
        P[t * 4 + K] = c ? b : a; P[t * 4 + J] = c ? a : b;
    }
    const int a = ix[K], b = ix[J];
    ix[K] = c ? b : a; ix[J] = c ? a : b;
}

__device__ void polar3(const double *mat, double *ortho) {
    double A[9], B[16];
    double fn = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) fn += mat[i] * mat[i];
    fn = sqrt(fn);
    if (fn == 0.0) {
#pragma unroll
        for (int i = 0; i < 9; i++) ortho[i] = 0.0;
        ortho[0] = ortho[4] = ortho[8] = 1.0;
        return;
    }
#pragma unroll
    for (int i = 0; i < 9; i++) A[i] = mat[i] / fn;
    double detB = 0, adj;
    adj = A[4] * A[8] - A[5] * A[7]; detB += adj * adj;
    adj = A[3] * A[8] - A[5] * A[6]; detB += adj * adj;
    adj = A[3] * A[7] - A[4] * A[6]; detB += adj * adj;
    adj = A[1] * A[8] - A[2] * A[7]; detB += adj * adj;
    adj = A[0] * A[8] - A[2] * A[6]; detB += adj * adj;
    adj = A[0] * A[7] - A[1] * A[6]; detB += adj * adj;
    adj = A[1] * A[5] - A[2] * A[4]; detB += adj * adj;
    adj = A[0] * A[5] - A[2] * A[3]; detB += adj * adj;
    adj = A[0] * A[4] - A[1] * A[3]; detB += adj * adj;
    detB = 1.0 - 4.0 * detB;
    B[0] = -A[0] - A[4] - A[8]; B[1] = -A[5] + A[7]; B[2] = -A[6] + A[2]; B[3] = -A[1] + A[3];
    B[4] = B[1]; B[5] = -A[0] + A[4] + A[8]; B[6] = -A[1] - A[3]; B[7] = -A[2] - A[6];
    B[8] = B[2]; B[9] = B[6]; B[10] = -A[4] + A[0] + A[8]; B[11] = -A[5] - A[7];
    B[12] = B[3]; B[13] = B[7]; B[14] = B[11]; B[15] = -A[8] + A[0] + A[4];

    double detA = 1.0, omega = 0.0;
    const double tau = 1e-4;
    const bool quick = detB <= 1.0 - tau;
    if (quick) {
        // determinant by LU with partial pivoting, on the rows as registers (the pivot row
        // and the two others in the reference's order; the same operations on the same values)
        const v3 R0 = mk(A[0], A[1], A[2]), R1 = mk(A[3], A[4], A[5]), R2 = mk(A[6], A[7], A[8]);
        int rowIndex = 0;
        double largest = fabs(A[0]);
        if (fabs(A[3]) > largest) { largest = fabs(A[3]); rowIndex = 1; }
        if (fabs(A[6]) > largest) rowIndex = 2;
        const v3 Pr = rowIndex == 0 ? R0 : (rowIndex == 1 ? R1 : R2);
        v3 Q0 = rowIndex == 1 ? R0 : R1;                 // row rp0
        v3 Q1 = rowIndex == 2 ? R0 : R2;                 // row rp1
        if (rowIndex != 0) detA = -detA;
        const double divisor = Pr.x;
        detA *= divisor;
        if (detA != 0.0) {
            const double lo0 = Q0.x / divisor, lo1 = Q1.x / divisor;
            Q0.y -= lo0 * Pr.y; Q0.z -= lo0 * Pr.z;
            Q1.y -= lo1 * Pr.y; Q1.z -= lo1 * Pr.z;
            v3 row, sub;
            if (fabs(Q0.y) > fabs(Q1.y)) { row = Q0; sub = Q1; }
            else { detA = -detA; row = Q1; sub = Q0; }
            const double div2 = row.y;
            detA *= div2;
            const double lower = sub.y / div2;
            detA *= (sub.z - lower * row.z);
        }
    } else {
        // determinant by LU with complete pivoting; omega = second pivot magnitude (on a local
        // copy: its runtime-indexed entries live in private memory, touched only on this path)
        double Aw[9];
#pragma unroll
        for (int i = 0; i < 9; i++) Aw[i] = A[i];
        int rp[3] = {0, 1, 2}, cp[3] = {0, 1, 2};
        for (int i = 0; i < 2; i++) {
            double largest = 0.0;
            int ri = i, ci = i;
            for (int j = i; j < 3; j++)
                for (int k = i; k < 3; k++) {
                    double en = fabs(Aw[rp[j] * 3 + cp[k]]);
                    if (en > largest) { largest = en; ri = j; ci = k; }
                }
            if (rp[i] != rp[ri]) { int t = rp[i]; rp[i] = rp[ri]; rp[ri] = t; detA = -detA; }
            if (cp[i] != cp[ci]) { int t = cp[i]; cp[i] = cp[ci]; cp[ci] = t; detA = -detA; }
            int row = rp[i], col = cp[i];
            double divisor = Aw[row * 3 + col];
            detA *= divisor;
            if (detA == 0.0) break;
            for (int j = i + 1; j < 3; j++) {
                int sr = rp[j];
                double lower = Aw[sr * 3 + col] / divisor;
                for (int k = i + 1; k < 3; k++) Aw[sr * 3 + cp[k]] -= lower * Aw[row * 3 + cp[k]];
            }
        }
        detA *= Aw[rp[2] * 3 + cp[2]];
        omega = fabs(Aw[rp[1] * 3 + cp[1]]);
    }
    const bool notpd = detA < 0.0;
    if (notpd) {
        detA = -detA;
#pragma unroll
        for (int i = 0; i < 16; i++) B[i] = -B[i];
    }
    const double inv3c = 0.333333333333333333333333333333333333;
    double dom;
    if (detB >= tau - inv3c) {
        double c = 8.0 * detA;
        double delta0 = 1.0 + 3.0 * detB;
        double delta1 = -1.0 + (27.0 / 16.0) * c * c + 9.0 * detB;
        double sd0 = sqrt(delta0);
        double alpha = delta1 / (delta0 * sd0);
        alpha = alpha < -1.0 ? -1.0 : (alpha > 1.0 ? 1.0 : alpha);
        double z = (4.0 * inv3c) * (1.0 + sd0 * cos(acos(alpha) * inv3c));
        double s = sqrt(z) * 0.5;
        double t = 4.0 - z + c / s;
        dom = s + 0.5 * sqrt(t > 0.0 ? t : 0.0);
    } else {
        dom = 1.7320508075688772935274463415058723669428052538;
        double old;
        double c = 8.0 * detA;
        int guard = 0;
        do {
            double p = dom * (dom * (dom * dom - 2.0) - c) + detB;
            double dp = dom * (4.0 * dom * dom - 4.0) - c;
            old = dom;
            dom -= p / dp;
        } while (old - dom > 1e-12 && ++guard < 100);
    }
    B[0] += dom; B[5] += dom; B[10] += dom; B[15] += dom;

    // LDL^T with symmetric pivoting of the 4x4 B (first two columns).  The pivots permute B
    // itself (rows and columns swapped, ix tracking which original index sits at each
    // position), so every entry is a register: position (i, j) holds the reference's
    // B[ix[i]][ix[j]] and sees the same operations in the same order.
    double L[6];
    int ix[4] = {0, 1, 2, 3};
    {
        int di = 0;
        double largest = B[0];
        if (B[5] > largest) { largest = B[5]; di = 1; }
        if (B[10] > largest) { largest = B[10]; di = 2; }
        if (B[15] > largest) { largest = B[15]; di = 3; }
        sym_swap4<0, 1>(B, ix, di == 1);
        sym_swap4<0, 2>(B, ix, di == 2);
        sym_swap4<0, 3>(B, ix, di == 3);
        L[0] = B[1] / largest;
        L[1] = B[2] / largest;
        L[2] = B[3] / largest;
        double od;
        B[5] -= L[0] * B[1];
        od = B[6] - L[0] * B[2]; B[6] = od; B[9] = od;
        od = B[7] - L[0] * B[3]; B[7] = od; B[13] = od;
        B[10] -= L[1] * B[2];
        od = B[11] - L[1] * B[3]; B[11] = od; B[14] = od;
        B[15] -= L[2] * B[3];
    }
    {
        int di = 1;
        double largest = B[5];
        if (B[10] > largest) { largest = B[10]; di = 2; }
        if (B[15] > largest) { largest = B[15]; di = 3; }
        sym_swap4<1, 2>(B, ix, di == 2);
        sym_swap4<1, 3>(B, ix, di == 3);
        const double l0 = L[0], l1 = L[1], l2 = L[2];
        L[0] = di == 2 ? l1 : (di == 3 ? l2 : l0);
        L[1] = di == 2 ? l0 : l1;
        L[2] = di == 3 ? l0 : l2;
        L[3] = B[6] / largest;
        L[4] = B[7] / largest;
        B[10] -= L[3] * B[6];
        const double od = B[11] - L[3] * B[7]; B[11] = od; B[14] = od;
        B[15] -= L[4] * B[7];
    }
    // the eigenvector by position (w[k] belongs at original index ix[k])
    double w[4];
    if (quick) {
        const bool sw = B[10] < B[15];
        sym_swap4<2, 3>(B, ix, sw);
        const double l1 = L[1], l2 = L[2], l3 = L[3], l4 = L[4];
        L[1] = sw ? l2 : l1; L[2] = sw ? l1 : l2;
        L[3] = sw ? l4 : l3; L[4] = sw ? l3 : l4;
        L[5] = B[11] / B[10];
        w[0] = L[0] * L[4] + L[1] * L[5] - L[0] * L[5] * L[3] - L[2];
        w[1] = L[5] * L[3] - L[4];
        w[2] = -L[5];
        w[3] = 1.0;
    } else {
        double b22 = B[10], b33 = B[15], b23 = B[11];
        double dd = b22 * b33 - b23 * b23;
        if (dd == 0.0) {
            if (b22 == 0.0 && b33 == 0.0 && b23 == 0.0) {
                w[0] = L[0] * L[4] - L[2];
                w[1] = -L[4];
                w[2] = 0.0; w[3] = 1.0;
            } else {
                double aa = -b23, bb = b22;
                double nrm = sqrt(aa * aa + bb * bb);
                aa /= nrm; bb /= nrm;
                w[3] = bb; w[2] = aa;
                w[1] = -L[4] * bb - L[3] * aa;
                w[0] = -L[2] * bb - L[1] * aa - L[0] * w[1];
            }
        } else {
            double ID0 = 1.0 / B[0], ID1 = 1.0 / B[5], ID2 = b33, ID3 = -b23, ID4 = b22;
            double IL0 = -L[0], IL1 = L[0] * L[3] - L[1], IL2 = L[0] * L[4] - L[2], IL3 = -L[3], IL4 = -L[4];
            if (omega > 6.607e-8) {
                double w0 = IL2, w1 = IL4, w2 = 0.0, w3 = 1.0;
                int nit = (int)ceil(15.0 / (16.86 + 2.0 * log10(omega)));
                for (int it = 0; it < nit; it++) {
                    double nrm = sqrt(w0 * w0 + w1 * w1 + w2 * w2 + w3 * w3);
                    w0 /= nrm; w1 /= nrm; w2 /= nrm; w3 /= nrm;
                    double t1 = w1;
                    w1 = IL0 * w0 + t1;
                    w2 = IL1 * w0 + IL3 * t1 + w2;
                    w3 = IL2 * w0 + IL4 * t1 + w3;
                    w0 *= ID0; w1 *= ID1;
                    double t2 = w2;
                    w2 = (ID2 * t2 + ID3 * w3) / dd;
                    w3 = (ID3 * t2 + ID4 * w3) / dd;
                    w0 = w0 + IL0 * w1 + IL1 * w2 + IL2 * w3;
                    w1 = w1 + IL3 * w2 + IL4 * w3;
                }
                w[0] = w0; w[1] = w1; w[2] = w2; w[3] = w3;
            } else {
                double vv[8] = {IL1, IL3, 1.0, 0.0, IL2, IL4, 0.0, 1.0};
                orthonormalize8(vv);
                for (int it = 0; it < 2; it++) {
                    double t1 = vv[1];
                    vv[1] = IL0 * vv[0] + t1;
                    vv[2] = IL1 * vv[0] + IL3 * t1 + vv[2];
                    vv[3] = IL2 * vv[0] + IL4 * t1 + vv[3];
                    double t5 = vv[5];
                    vv[5] = IL0 * vv[4] + t5;
                    vv[6] = IL1 * vv[4] + IL3 * t5 + vv[6];
                    vv[7] = IL2 * vv[4] + IL4 * t5 + vv[7];
                    vv[0] *= ID0; vv[1] *= ID1;
                    double t2 = vv[2];
                    vv[2] = (ID2 * t2 + ID3 * vv[3]) / dd;
                    vv[3] = (ID3 * t2 + ID4 * vv[3]) / dd;
                    // the reference scales the second vector's entry 1 by ID2 (numerMethod.h:928)
                    vv[4] *= ID0; vv[5] *= ID2;
                    double t6 = vv[6];
                    vv[6] = (ID2 * t6 + ID3 * vv[7]) / dd;
                    vv[7] = (ID3 * t6 + ID4 * vv[7]) / dd;
                    vv[0] = vv[0] + IL0 * vv[1] + IL1 * vv[2] + IL2 * vv[3];
                    vv[1] = vv[1] + IL3 * vv[2] + IL4 * vv[3];
                    vv[4] = vv[4] + IL0 * vv[5] + IL1 * vv[6] + IL2 * vv[7];
                    vv[5] = vv[5] + IL3 * vv[6] + IL4 * vv[7];
                }
                orthonormalize8(vv);
                double T0[8], T1[8];
                T0[0] = vv[0] + L[0] * vv[1] + L[1] * vv[2] + L[2] * vv[3];
                T0[1] = vv[1] + L[3] * vv[2] + L[4] * vv[3];
                T0[2] = vv[2]; T0[3] = vv[3];
                T0[4] = vv[4] + L[0] * vv[5] + L[1] * vv[6] + L[2] * vv[7];
                T0[5] = vv[5] + L[3] * vv[6] + L[4] * vv[7];
                T0[6] = vv[6]; T0[7] = vv[7];
                double d00 = B[0], d11 = B[5];
                T1[0] = T0[0] * d00; T1[1] = T0[1] * d11;
                T1[2] = T0[2] * b22 + T0[3] * b23; T1[3] = T0[2] * b23 + T0[3] * b33;
                T1[4] = T0[4] * d00; T1[5] = T0[5] * d11;
                T1[6] = T0[6] * b22 + T0[7] * b23; T1[7] = T0[6] * b23 + T0[7] * b33;
                double BB0 = T1[0] * T0[0] + T1[1] * T0[1] + T1[2] * T0[2] + T1[3] * T0[3];
                double BB1 = T1[0] * T0[4] + T1[1] * T0[5] + T1[2] * T0[6] + T1[3] * T0[7];
                double BB2 = T1[4] * T0[4] + T1[5] * T0[5] + T1[6] * T0[6] + T1[7] * T0[7];
                if (fabs(BB1) > 1e-15) {
                    double rr = (BB0 - BB2) * 0.5;
                    double w0 = rr - sqrt(rr * rr + BB1 * BB1);
                    for (int k = 0; k < 4; k++) w[k] = vv[k] * w0 + vv[4 + k] * BB1;
                } else if (BB0 < BB2) {
                    for (int k = 0; k < 4; k++) w[k] = vv[k];
                } else {
                    for (int k = 0; k < 4; k++) w[k] = vv[4 + k];
                }
            }
        }
    }
    double v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = ix[0] == j ? w[0] : (ix[1] == j ? w[1] : (ix[2] == j ? w[2] : w[3]));
    double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
    v[0] /= nrm; v[1] /= nrm; v[2] /= nrm; v[3] /= nrm;
    double v11 = v[1] * v[1], v22 = v[2] * v[2], v33 = v[3] * v[3];
    double v12 = v[1] * v[2], v03 = v[0] * v[3], v13 = v[1] * v[3];
    double v02 = v[0] * v[2], v23 = v[2] * v[3], v01 = v[0] * v[1];
    double sg = notpd ? -1.0 : 1.0;
    // ortho = transpose of the quaternion matrix Q (numerMethod.h:998-1007)
    ortho[0] = sg * (1.0 - 2.0 * (v22 + v33));
    ortho[3] = sg * (2.0 * (v12 - v03));
    ortho[6] = sg * (2.0 * (v13 + v02));
    ortho[1] = sg * (2.0 * (v12 + v03));
    ortho[4] = sg * (1.0 - 2.0 * (v11 + v33));
    ortho[7] = sg * (2.0 * (v23 - v01));
    ortho[2] = sg * (2.0 * (v13 - v02));
    ortho[5] = sg * (2.0 * (v23 + v01));
    ortho[8] = sg * (1.0 - 2.0 * (v11 + v22));
}

// Closed-form symmetric 3x3 eigenpairs; tri = [a00 a01 a02 a11 a12 a22]; vectors as rows.
__device__ void eigsym3(const double *tri, double *ev, double *vec) {
    const double pi = 3.14159265358979323846;
    const double ave = (tri[0] + tri[3] + tri[5]) / 3.0;
    const double d0 = tri[0] - ave, d1 = tri[3] - ave, d2 = tri[5] - ave;
    const double inv = (d0 * d0 + d1 * d1 + d2 * d2) * 0.5 + tri[1] * tri[1] + tri[2] * tri[2] + tri[4] * tri[4];
    const double det = d0 * (d1 * d2 - tri[4] * tri[4]) + tri[1] * (tri[4] * tri[2] - tri[1] * d2) + tri[2] * (tri[1] * tri[4] - d1 * tri[2]);
    const double sq3 = 1.7320508075688772935274463415058723669428052538;
    const double st = sqrt(inv);
    const double ac = (3.0 * 0.5 * sq3) * ((det / inv) / st);
    if (isnan(ac) || isinf(ac)) {
        ev[0] = ev[1] = ev[2] = ave;
        for (int i = 0; i < 9; i++) vec[i] = 0.0;
        vec[0] = vec[4] = vec[8] = 1.0;
        return;
    }
    double cl = ac < -1.0 ? -1.0 : (ac > 1.0 ? 1.0 : ac);
    double angle = acos(cl) / 3.0;
    double dist = (2.0 / sq3) * st * (angle < pi / 6.0 ? cos(angle) : cos(angle + (2.0 / 3.0) * pi));
    ev[0] = dist + ave;
    v3 r0 = mk(d0 - dist, tri[1], tri[2]);
    v3 r1 = mk(tri[1], d1 - dist, tri[4]);
    v3 r2 = mk(tri[2], tri[4], d2 - dist);
    double l0 = sqrt(dot(r0, r0)), l1 = sqrt(dot(r1, r1)), l2 = sqrt(dot(r2, r2));
    v3 s0, t1, t2;
    if (l0 > l1) {
        if (l0 > l2) { s0 = vdiv(r0, l0); t1 = r1 - s0 * dot(s0, r1); t2 = r2 - s0 * dot(s0, r2); }
        else { s0 = vdiv(r2, l2); t1 = r0 - s0 * dot(s0, r0); t2 = r1 - s0 * dot(s0, r1); }
    } else {
        if (l1 > l2) { s0 = vdiv(r1, l1); t1 = r0 - s0 * dot(s0, r0); t2 = r2 - s0 * dot(s0, r2); }
        else { s0 = vdiv(r2, l2); t1 = r0 - s0 * dot(s0, r0); t2 = r1 - s0 * dot(s0, r1); }
    }
    double q1 = dot(t1, t1), q2 = dot(t2, t2);
    v3 s1 = q1 > q2 ? vdiv(t1, sqrt(q1)) : vdiv(t2, sqrt(q2));
    v3 e0 = cross(s0, s1);
    vec[0] = e0.x; vec[1] = e0.y; vec[2] = e0.z;
    double m[9] = {d0, tri[1], tri[2], tri[1], d1, tri[4], tri[2], tri[4], d2};
    v3 tr0 = mulv(m, s0), tr1 = mulv(m, s1);
    double a11 = dot(s0, tr0), a12 = dot(s0, tr1), a22 = dot(s1, tr1);
    double disc = sqrt((a11 - a22) * (a11 - a22) + 4.0 * a12 * a12);
    double e1v = ((a11 + a22) - (a11 > a22 ? disc : -disc)) / 2.0;
    double e2v = a11 + a22 - e1v;
    ev[1] = e1v + ave;
    ev[2] = e2v + ave;
    v3 u0 = cross(tr0 - s0 * e1v, e0);
    v3 u1 = cross(tr1 - s1 * e1v, e0);
    double u0l = dot(u0, u0), u1l = dot(u1, u1);
    v3 e1 = (u0l != 0.0) ? (u0l > u1l ? vdiv(u0, sqrt(u0l)) : vdiv(u1, sqrt(u1l))) : s0;
    vec[3] = e1.x; vec[4] = e1.y; vec[5] = e1.z;
    v3 e2 = cross(e0, e1);
    vec[6] = e2.x; vec[7] = e2.y; vec[8] = e2.z;
}

__device__ __forceinline__ void coord_system(v3 v0, v3 &v1, v3 &v2) {
    if (fabs(v0.x) > fabs(v0.y)) {
        double il = 1.0 / sqrt(v0.x * v0.x + v0.z * v0.z);
        v1 = mk(-v0.z * il, 0.0, v0.x * il);
    } else {
        double il = 1.0 / sqrt(v0.y * v0.y + v0.z * v0.z);
        v1 = mk(0.0, -v0.z * il, v0.y * il);
    }
    v2 = cross(v0, v1);
}

// Diagonalised deformation gradient with inversion handling and singular-value clamp.
// Outputs sigma (3), U (row-major, U = UT^T) and VT (row-major).
__device__ void neo_diag(const double *F, double thr, double *sig, double *U, double *VT) {
    double tri[6], ev[3], UT[9];
    tri[0] = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    tri[1] = F[3] * F[0] + F[4] * F[1] + F[5] * F[2];
    tri[2] = F[6] * F[0] + F[7] * F[1] + F[8] * F[2];
    tri[3] = F[3] * F[3] + F[4] * F[4] + F[5] * F[5];
    tri[4] = F[6] * F[3] + F[7] * F[4] + F[8] * F[5];
    tri[5] = F[6] * F[6] + F[7] * F[7] + F[8] * F[8];
    eigsym3(tri, ev, UT);
    if (det3(UT) < 0.0) { UT[0] = -UT[0]; UT[1] = -UT[1]; UT[2] = -UT[2]; }
    for (int i = 0; i < 3; i++) ev[i] = ev[i] > 0.0 ? sqrt(ev[i]) : 0.0;
    mul3(UT, F, VT);
    const double eps = 1e-8;
    int cond = (ev[0] < eps) + ((ev[1] < eps) << 1) + ((ev[2] < eps) << 2);
    switch (cond) {
    case 0: {
        for (int i = 0; i < 3; i++) { double iv = 1.0 / ev[i]; VT[i * 3] *= iv; VT[i * 3 + 1] *= iv; VT[i * 3 + 2] *= iv; }
        if (det3(VT) < 0.0) {
            const int sm = ev[0] < ev[1] ? (ev[0] < ev[2] ? 0 : 2) : (ev[1] < ev[2] ? 1 : 2);
            // row sm negated, by selects (a runtime index would put VT in private memory)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const bool ng = sm == r;
                VT[r * 3] = ng ? -VT[r * 3] : VT[r * 3];
                VT[r * 3 + 1] = ng ? -VT[r * 3 + 1] : VT[r * 3 + 1];
                VT[r * 3 + 2] = ng ? -VT[r * 3 + 2] : VT[r * 3 + 2];
                ev[r] = ng ? -ev[r] : ev[r];
            }
        }
        break;
    }
    case 1: {
        double i1 = 1.0 / ev[1], i2 = 1.0 / ev[2];
        for (int j = 0; j < 3; j++) { VT[3 + j] *= i1; VT[6 + j] *= i2; }
        v3 an = cross(mk(VT[3], VT[4], VT[5]), mk(VT[6], VT[7], VT[8]));
        VT[0] = an.x; VT[1] = an.y; VT[2] = an.z;
        break;
    }
    case 2: {
        double i0 = 1.0 / ev[0], i2 = 1.0 / ev[2];
        for (int j = 0; j < 3; j++) { VT[j] *= i0; VT[6 + j] *= i2; }
        v3 an = cross(mk(VT[6], VT[7], VT[8]), mk(VT[0], VT[1], VT[2]));
        VT[3] = an.x; VT[4] = an.y; VT[5] = an.z;
        break;
    }
    case 3: {
        double iv = 1.0 / ev[2];
        VT[6] *= iv; VT[7] *= iv; VT[8] *= iv;
        v3 a1, a2;
        coord_system(mk(VT[6], VT[7], VT[8]), a1, a2);
        VT[0] = a1.x; VT[1] = a1.y; VT[2] = a1.z; VT[3] = a2.x; VT[4] = a2.y; VT[5] = a2.z;
        break;
    }
    case 4: {
        double i0 = 1.0 / ev[0], i1 = 1.0 / ev[1];
        for (int j = 0; j < 3; j++) { VT[j] *= i0; VT[3 + j] *= i1; }
        v3 an = cross(mk(VT[0], VT[1], VT[2]), mk(VT[3], VT[4], VT[5]));
        VT[6] = an.x; VT[7] = an.y; VT[8] = an.z;
        break;
    }
    case 5: {
        double iv = 1.0 / ev[1];
        VT[3] *= iv; VT[4] *= iv; VT[5] *= iv;
        v3 a1, a2;
        coord_system(mk(VT[3], VT[4], VT[5]), a1, a2);
        VT[6] = a1.x; VT[7] = a1.y; VT[8] = a1.z; VT[0] = a2.x; VT[1] = a2.y; VT[2] = a2.z;
        break;
    }
    case 6: {
        double iv = 1.0 / ev[0];
        VT[0] *= iv; VT[1] *= iv; VT[2] *= iv;
        v3 a1, a2;
        coord_system(mk(VT[0], VT[1], VT[2]), a1, a2);
        VT[3] = a1.x; VT[4] = a1.y; VT[5] = a1.z; VT[6] = a2.x; VT[7] = a2.y; VT[8] = a2.z;
        break;
    }
    default: {
        for (int i = 0; i < 9; i++) VT[i] = 0.0;
        VT[0] = VT[4] = VT[8] = 1.0;
        break;
    }
    }
    for (int i = 0; i < 3; i++) sig[i] = ev[i] > thr ? ev[i] : thr;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) U[i * 3 + j] = UT[j * 3 + i];
}

__device__ __forceinline__ void psd2(double &dg, double &od) {
    double e0 = dg + od, e1 = dg - od;
    int c = (e0 < 0.0) + ((e1 < 0.0) << 1);
    if (c == 1) { e1 *= 0.5; dg = e1; od = -e1; }
    else if (c == 2) { e0 *= 0.5; dg = e0; od = e0; }
    else if (c == 3) { dg = 0.0; od = 0.0; }
}

__device__ __forceinline__ void psd3(double *m) {
    double QT[9], ev[3];
    eigsym3(m, ev, QT);
    bool neg = false;
    for (int i = 0; i < 3; i++) if (ev[i] < 0.0) { neg = true; ev[i] = 0.0; }
    if (neg) {
        // A = Q diag(ev) Q^T with Q = QT^T
        double A[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                A[i * 3 + j] = QT[0 * 3 + i] * ev[0] * QT[0 * 3 + j] + QT[1 * 3 + i] * ev[1] * QT[1 * 3 + j] + QT[2 * 3 + i] * ev[2] * QT[2 * 3 + j];
        m[0] = A[0]; m[1] = A[1]; m[2] = A[2]; m[3] = A[4]; m[4] = A[5]; m[5] = A[8];
    }
}

// Sparse dP/dF of the invertible neo-Hookean model in the diagonal frame:
// tm = PSD-projected (i,i,k,k) block, al/be = (i,j,i,j)/(i,j,j,i) pairs for (01,02,12).
// Also returns P = U diag(...) VT.
__device__ void neo_stress(const double *F, double lambda, double mu, double thr, double *U, double *VT,
                           double *P, double *tm, double *al, double *be) {
    double s[3];
    neo_diag(F, thr, s, U, VT);
    double s0 = s[0] * s[0], s1 = s[1] * s[1], s2 = s[2] * s[2];
    double I3 = s0 * s1 * s2;
    double lnI3 = log(I3);
    double g0 = 0.5 * mu, g1 = 0.0, g2 = (-0.5 * mu + 0.25 * lambda * lnI3) / I3;
    double h5 = (0.5 * mu + 0.25 * lambda * (1.0 - lnI3)) / (I3 * I3);
    double de[3];
    for (int i = 0; i < 3; i++) de[i] = 2.0 * g0 * s[i] + 4.0 * g1 * s[i] * s[i] * s[i] + 2.0 * g2 * I3 / s[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            P[i * 3 + j] = U[i * 3 + 0] * VT[0 * 3 + j] * de[0] + U[i * 3 + 1] * VT[1 * 3 + j] * de[1] + U[i * 3 + 2] * VT[2 * 3 + j] * de[2];
    const double ds[3] = {s0, s1, s2};
    const double dc[3] = {s[0] * s[1], s[0] * s[2], s[1] * s[2]};
    double a00 = 2.0 * g0 + 8.0 * ds[0] * g1, a11 = 2.0 * g0 + 8.0 * ds[1] * g1, a22 = 2.0 * g0 + 8.0 * ds[2] * g1;
    double a01 = 2.0 * g0 + 4.0 * (ds[0] + ds[1]) * g1, a02 = 2.0 * g0 + 4.0 * (ds[0] + ds[2]) * g1, a12 = 2.0 * g0 + 4.0 * (ds[1] + ds[2]) * g1;
    double d00 = I3 * g2 / ds[0], d01 = I3 * g2 / dc[0], d02 = I3 * g2 / dc[1];
    double d11 = I3 * g2 / ds[1], d12 = I3 * g2 / dc[2], d22 = I3 * g2 / ds[2];
    double b00 = 4.0 * ds[0] * g1 - 2.0 * d00, b01 = 4.0 * dc[0] * g1 - 2.0 * d01, b02 = 4.0 * dc[1] * g1 - 2.0 * d02;
    double b11 = 4.0 * ds[1] * g1 - 2.0 * d11, b12 = 4.0 * dc[2] * g1 - 2.0 * d12, b22 = 4.0 * ds[2] * g1 - 2.0 * d22;
    // energy Hessian has only its (I3, I3) entry: vec_i^T H vec_j = h5 * (2 I3 / s_i) (2 I3 / s_j)
    double w0 = 2.0 * I3 / s[0], w1 = 2.0 * I3 / s[1], w2 = 2.0 * I3 / s[2];
    double c00 = w0 * h5 * w0 + 4.0 * d00, c01 = w1 * h5 * w0 + 4.0 * d01, c02 = w2 * h5 * w0 + 4.0 * d02;
    double c11 = w1 * h5 * w1 + 4.0 * d11, c12 = w2 * h5 * w1 + 4.0 * d12, c22 = w2 * h5 * w2 + 4.0 * d22;
    tm[0] = a00 + b00 + c00; tm[1] = c01; tm[2] = c02;
    tm[3] = a11 + b11 + c11; tm[4] = c12; tm[5] = a22 + b22 + c22;
    psd3(tm);
    psd2(a01, b01); psd2(a02, b02); psd2(a12, b12);
    al[0] = a01; al[1] = a02; al[2] = a12;
    be[0] = b01; be[1] = b02; be[2] = b12;
}

}  // namespace odk
