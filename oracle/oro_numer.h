// ORACLE — test infrastructure only. Never linked into the product library.
//
// FP64 CPU restatement of the reference's dense 3x3 numerics.  Every routine
// follows the reference file:line named in its comment, branch for branch, so
// that the GPU kernels can be checked against it.  Nothing here is used by the
// HIP path; only tests, smoke() and the bench's cpu_baseline leg load it.
#pragma once
#include <cmath>
#include <cstring>
#include <algorithm>
#include <limits>
#include <utility>

namespace oro {
typedef double Scalar;

// core/latool.h:12-117 (VectorBase): +, -, scalar *, dot (*), cross (%), / as multiply by reciprocal
struct Vec3 {
    Scalar x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(Scalar a, Scalar b, Scalar c) : x(a), y(b), z(c) {}
    Scalar operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Scalar &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    Vec3 operator+(const Vec3 &v) const { return Vec3(x + v.x, y + v.y, z + v.z); }
    Vec3 operator-(const Vec3 &v) const { return Vec3(x - v.x, y - v.y, z - v.z); }
    Vec3 operator-() const { return Vec3(-x, -y, -z); }
    Vec3 operator*(Scalar f) const { return Vec3(x * f, y * f, z * f); }
    friend Vec3 operator*(Scalar f, const Vec3 &v) { return v * f; }
    Scalar operator*(const Vec3 &v) const { return x * v.x + y * v.y + z * v.z; }
    Vec3 operator/(Scalar f) const { Scalar inv = Scalar(1.0) / f; return Vec3(inv * x, inv * y, inv * z); }
    Vec3 operator%(const Vec3 &v) const { return Vec3(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x); }
    Scalar length2() const { return x * x + y * y + z * z; }
    Scalar length() const { return std::sqrt(length2()); }
};

// core/latool.h:119-220 (Tensor2, row-major m[3][3])
struct Ten3 {
    Scalar m[3][3];
    Ten3() { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m[i][j] = 0; }
    explicit Ten3(const Scalar *p) { std::memcpy(m, p, 9 * sizeof(Scalar)); }
    Ten3(Scalar a00, Scalar a01, Scalar a02, Scalar a10, Scalar a11, Scalar a12, Scalar a20, Scalar a21, Scalar a22) {
        m[0][0] = a00; m[0][1] = a01; m[0][2] = a02;
        m[1][0] = a10; m[1][1] = a11; m[1][2] = a12;
        m[2][0] = a20; m[2][1] = a21; m[2][2] = a22;
    }
    Scalar &operator()(int r, int c) { return m[r][c]; }
    Scalar operator()(int r, int c) const { return m[r][c]; }
    Ten3 operator*(const Ten3 &b) const {
        Ten3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                r.m[i][j] = m[i][0] * b.m[0][j] + m[i][1] * b.m[1][j] + m[i][2] * b.m[2][j];
        return r;
    }
    Vec3 operator*(const Vec3 &v) const {
        return Vec3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
                    m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                    m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
    }
    // latool.h:184-188
    Scalar Determinant() const {
        return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
             - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
             + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    }
};

// latool.h:189-208 (adjugate times 1/det)
inline Ten3 Inverse(const Ten3 &t) {
    Ten3 r;
    Scalar det = t.Determinant();
    Scalar invDet = 1.f / det;
    r.m[0][0] = (t.m[1][1] * t.m[2][2] - t.m[1][2] * t.m[2][1]) * invDet;
    r.m[0][1] = (t.m[0][2] * t.m[2][1] - t.m[0][1] * t.m[2][2]) * invDet;
    r.m[0][2] = (t.m[0][1] * t.m[1][2] - t.m[0][2] * t.m[1][1]) * invDet;
    r.m[1][0] = (t.m[1][2] * t.m[2][0] - t.m[1][0] * t.m[2][2]) * invDet;
    r.m[1][1] = (t.m[0][0] * t.m[2][2] - t.m[0][2] * t.m[2][0]) * invDet;
    r.m[1][2] = (t.m[0][2] * t.m[1][0] - t.m[0][0] * t.m[1][2]) * invDet;
    r.m[2][0] = (t.m[1][0] * t.m[2][1] - t.m[1][1] * t.m[2][0]) * invDet;
    r.m[2][1] = (t.m[0][1] * t.m[2][0] - t.m[0][0] * t.m[2][1]) * invDet;
    r.m[2][2] = (t.m[0][0] * t.m[1][1] - t.m[0][1] * t.m[1][0]) * invDet;
    return r;
}

inline Ten3 Transpose(const Ten3 &a) {
    Ten3 r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i];
    return r;
}

inline Scalar Clamp(Scalar v, Scalar lo, Scalar hi) { return v < lo ? lo : (v > hi ? hi : v); }  // oder.h:121

// latool.h:542-552
inline void coordinateSystem(const Vec3 &v0, Vec3 &v1, Vec3 &v2) {
    if (std::fabs(v0.x) > std::fabs(v0.y)) {
        Scalar invlen = Scalar(1.0) / std::sqrt(v0.x * v0.x + v0.z * v0.z);
        v1 = Vec3(-v0.z * invlen, Scalar(0.0), v0.x * invlen);
    } else {
        Scalar invlen = 1.f / std::sqrt(v0.y * v0.y + v0.z * v0.z);
        v1 = Vec3(Scalar(0.0), -v0.z * invlen, v0.y * invlen);
    }
    v2 = v0 % v1;
}

// numerMethod.h:327-429 — analytic symmetric 3x3 eigen-decomposition.
// triMat packed upper: [a00 a01 a02 a11 a12 a22]; eigenvectors written as rows.
inline void eigenSym3x3(const Scalar *triMat, Scalar *eigenvalues, Scalar *eigenvectors) {
    const Scalar pi = Scalar(M_PI);
    const Scalar aveTrace = (triMat[0] + triMat[3] + triMat[5]) / Scalar(3.0);
    const Scalar dd0 = triMat[0] - aveTrace;
    const Scalar dd1 = triMat[3] - aveTrace;
    const Scalar dd2 = triMat[5] - aveTrace;
    const Scalar invariant = (dd0 * dd0 + dd1 * dd1 + dd2 * dd2) * Scalar(0.5) +
        triMat[1] * triMat[1] + triMat[2] * triMat[2] + triMat[4] * triMat[4];
    const Scalar det = dd0 * (dd1 * dd2 - triMat[4] * triMat[4]) +
        triMat[1] * (triMat[4] * triMat[2] - triMat[1] * dd2) +
        triMat[2] * (triMat[1] * triMat[4] - dd1 * triMat[2]);
    const Scalar squareRoot3 = Scalar(1.7320508075688772935274463415058723669428052538103806280558069794519330169088);
    const Scalar sqrtTerm = std::sqrt(invariant);
    const Scalar angleCos = (Scalar(3) * Scalar(0.5) * squareRoot3) * ((det / invariant) / sqrtTerm);
    if (std::isnan(angleCos) || std::isinf(angleCos)) {
        eigenvalues[0] = eigenvalues[1] = eigenvalues[2] = aveTrace;
        for (int i = 0; i < 9; i++) eigenvectors[i] = 0;
        eigenvectors[0] = eigenvectors[4] = eigenvectors[8] = Scalar(1);
        return;
    }
    Scalar angle = std::acos(Clamp(angleCos, Scalar(-1), Scalar(1))) / Scalar(3);
    Scalar distinct = (Scalar(2) / squareRoot3) * sqrtTerm *
        (angle < pi / Scalar(6) ? std::cos(angle) : std::cos(angle + (Scalar(2) / Scalar(3)) * pi));
    eigenvalues[0] = distinct + aveTrace;

    const Vec3 r0(dd0 - distinct, triMat[1], triMat[2]);
    const Vec3 r1(triMat[1], dd1 - distinct, triMat[4]);
    const Vec3 r2(triMat[2], triMat[4], dd2 - distinct);
    const Scalar r0Len = r0.length(), r1Len = r1.length(), r2Len = r2.length();
    Vec3 s0, t1, t2;
    if (r0Len > r1Len) {
        if (r0Len > r2Len) { s0 = r0 / r0Len; t1 = r1 - (s0 * r1) * s0; t2 = r2 - (s0 * r2) * s0; }
        else { s0 = r2 / r2Len; t1 = r0 - (s0 * r0) * s0; t2 = r1 - (s0 * r1) * s0; }
    } else {
        if (r1Len > r2Len) { s0 = r1 / r1Len; t1 = r0 - (s0 * r0) * s0; t2 = r2 - (s0 * r2) * s0; }
        else { s0 = r2 / r2Len; t1 = r0 - (s0 * r0) * s0; t2 = r1 - (s0 * r1) * s0; }
    }
    const Scalar t1Len2 = t1.length2(), t2Len2 = t2.length2();
    const Vec3 s1 = t1Len2 > t2Len2 ? t1 / std::sqrt(t1Len2) : t2 / std::sqrt(t2Len2);
    Vec3 ev0 = s0 % s1;
    eigenvectors[0] = ev0.x; eigenvectors[1] = ev0.y; eigenvectors[2] = ev0.z;

    Ten3 mat(dd0, triMat[1], triMat[2], triMat[1], dd1, triMat[4], triMat[2], triMat[4], dd2);
    const Vec3 trans0 = mat * s0;
    const Vec3 trans1 = mat * s1;
    const Scalar a11 = s0 * trans0, a12 = s0 * trans1, a22 = s1 * trans1;
    const Scalar disc = std::sqrt((a11 - a22) * (a11 - a22) + Scalar(4) * a12 * a12);
    const Scalar ev1v = ((a11 + a22) - (a11 > a22 ? disc : -disc)) / Scalar(2);
    const Scalar ev2v = a11 + a22 - ev1v;
    eigenvalues[1] = ev1v + aveTrace;
    eigenvalues[2] = ev2v + aveTrace;

    Vec3 ev1;
    const Vec3 u0 = (trans0 - ev1v * s0) % ev0;
    const Vec3 u1 = (trans1 - ev1v * s1) % ev0;
    const Scalar u0Len2 = u0.length2(), u1Len2 = u1.length2();
    // numerMethod.h:422 tests u0Len2 twice (reference quirk, kept)
    if (u0Len2 != Scalar(0) && u0Len2 != Scalar(0)) ev1 = u0Len2 > u1Len2 ? u0 / std::sqrt(u0Len2) : u1 / std::sqrt(u1Len2);
    else ev1 = s0;
    eigenvectors[3] = ev1.x; eigenvectors[4] = ev1.y; eigenvectors[5] = ev1.z;
    Vec3 ev2 = ev0 % ev1;
    eigenvectors[6] = ev2.x; eigenvectors[7] = ev2.y; eigenvectors[8] = ev2.z;
}

// numerMethod.h:431-507 — Givens QR with proper-rotation sign convention
inline void properQRDecompostion3x3(const Scalar *mat, Scalar *ortho, Scalar *upperTri) {
    auto givens = [](Scalar a, Scalar b, Scalar &c, Scalar &s, Scalar &r) {
        if (b == Scalar(0)) { c = a < Scalar(0) ? Scalar(-1) : Scalar(1); s = Scalar(0); r = std::fabs(a); }
        else if (a == Scalar(0)) { c = Scalar(0); s = b < Scalar(0) ? Scalar(1) : Scalar(-1); r = std::fabs(b); }
        else if (std::fabs(b) > std::fabs(a)) {
            Scalar tau = -a / b; Scalar u = std::sqrt(Scalar(1) + tau * tau);
            if (b < Scalar(0)) u = -u;
            s = -Scalar(1) / u; c = s * tau; r = b * u;
        } else {
            Scalar tau = -b / a; Scalar u = std::sqrt(Scalar(1) + tau * tau);
            if (a < Scalar(0)) u = -u;
            c = Scalar(1) / u; s = c * tau; r = a * u;
        }
    };
    Scalar Q[9], R[9];
    std::memcpy(R, mat, sizeof(Scalar) * 9);
    for (int i = 0; i < 9; i++) Q[i] = 0;
    Scalar c = 0, s = 0, r = 0;
    givens(R[3], R[6], c, s, r);
    Q[4] = Q[8] = c; Q[5] = s; Q[7] = -s;
    R[3] = r; R[6] = Scalar(0);
    for (int i = 0; i < 2; i++) {
        Scalar p = c * R[4 + i] - s * R[7 + i];
        Scalar q = s * R[4 + i] + c * R[7 + i];
        R[4 + i] = p; R[7 + i] = q;
    }
    givens(R[0], R[3], c, s, r);
    R[0] = r; R[3] = Scalar(0);
    for (int i = 0; i < 2; i++) {
        Scalar p = c * R[1 + i] - s * R[4 + i];
        Scalar q = s * R[1 + i] + c * R[4 + i];
        R[1 + i] = p; R[4 + i] = q;
    }
    Q[0] = c; Q[1] = s;
    Q[3] = Q[4] * -s; Q[4] *= c;
    Q[6] = Q[7] * -s; Q[7] *= c;
    givens(R[4], R[7], c, s, r);
    R[4] = r; R[7] = Scalar(0);
    Scalar pp = c * R[5] - s * R[8];
    Scalar qq = s * R[5] + c * R[8];
    R[5] = pp; R[8] = qq;
    for (int i = 0; i < 3; i++) {
        Scalar p = c * Q[i * 3 + 1] - s * Q[i * 3 + 2];
        Scalar q = s * Q[i * 3 + 1] + c * Q[i * 3 + 2];
        Q[i * 3 + 1] = p; Q[i * 3 + 2] = q;
    }
    std::memcpy(ortho, Q, sizeof(Scalar) * 9);
    if (upperTri) std::memcpy(upperTri, R, sizeof(Scalar) * 9);
}

// numerMethod.h:509-568 — two Householder reflections orthonormalising a 4x2 block
inline void Orthonormalize(Scalar mat[8]) {
    const Scalar tol = std::numeric_limits<Scalar>::min();
    Scalar beta0 = 0, beta1 = 0;
    Scalar v0[3], v1[2];
    v0[0] = mat[1]; v0[1] = mat[2]; v0[2] = mat[3];
    Scalar norm2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    if (norm2 <= tol) { v0[0] = v0[1] = v0[2] = 0; }
    else {
        Scalar c0 = mat[0];
        Scalar sigma = std::sqrt(c0 * c0 + norm2);
        if (c0 >= Scalar(0)) sigma = -sigma;
        Scalar factor = c0 - sigma;
        v0[0] /= factor; v0[1] /= factor; v0[2] /= factor;
        beta0 = (sigma - c0) / sigma;
    }
    Scalar product = mat[4] + v0[0] * mat[5] + v0[1] * mat[6] + v0[2] * mat[7];
    Scalar c0 = mat[5] - product * beta0 * v0[0];
    v1[0] = mat[6] - product * beta0 * v0[1];
    v1[1] = mat[7] - product * beta0 * v0[2];
    norm2 = v1[0] * v1[0] + v1[1] * v1[1];
    if (norm2 <= tol) { v1[0] = v1[1] = 0; }
    else {
        Scalar sigma = std::sqrt(c0 * c0 + norm2);
        if (c0 >= Scalar(0)) sigma = -sigma;
        Scalar factor = c0 - sigma;
        v1[0] /= factor; v1[1] /= factor;
        beta1 = (sigma - c0) / sigma;
    }
    mat[0] = Scalar(1) - beta0; mat[1] = -beta0 * v0[0]; mat[2] = -beta0 * v0[1]; mat[3] = -beta0 * v0[2];
    Scalar subMat[9], lhs[3];
    subMat[0] = Scalar(1) - beta0 * v0[0] * v0[0];
    subMat[1] = -beta0 * v0[0] * v0[1]; subMat[3] = subMat[1];
    subMat[2] = -beta0 * v0[0] * v0[2]; subMat[6] = subMat[2];
    subMat[4] = Scalar(1) - beta0 * v0[1] * v0[1];
    subMat[5] = -beta0 * v0[1] * v0[2]; subMat[7] = subMat[5];
    subMat[8] = Scalar(1) - beta0 * v0[2] * v0[2];
    lhs[0] = Scalar(1) - beta1; lhs[1] = -beta1 * v1[0]; lhs[2] = -beta1 * v1[1];
    mat[4] = mat[1] * lhs[0] + mat[2] * lhs[1] + mat[3] * lhs[2];
    mat[5] = subMat[0] * lhs[0] + subMat[1] * lhs[1] + subMat[2] * lhs[2];
    mat[6] = subMat[3] * lhs[0] + subMat[4] * lhs[1] + subMat[5] * lhs[2];
    mat[7] = subMat[6] * lhs[0] + subMat[7] * lhs[1] + subMat[8] * lhs[2];
}

// numerMethod.h:570-1016 — polar decomposition through the dominant eigenvector
// of the 4x4 quaternion matrix B (Higham & Noferini).  ortho receives the
// transposed rotation, sspd (optional) the symmetric factor, as the reference.  Pinned
// against numpy's SVD polar factor and special cases; parity unpinned: which branch is
// taken for inputs exactly at its thresholds (numerMethod.h:603-604, 715, 873-877).
inline void polarDecompostion3x3(const Scalar *mat, Scalar *ortho, Scalar *sspd) {
    Scalar A[9], B[16];
    Scalar fnorm = 0;
    for (int i = 0; i < 9; i++) fnorm += mat[i] * mat[i];
    fnorm = std::sqrt(fnorm);
    if (fnorm == Scalar(0)) {
        for (int i = 0; i < 9; i++) ortho[i] = 0;
        if (sspd) for (int i = 0; i < 9; i++) sspd[i] = 0;
        ortho[0] = ortho[4] = ortho[8] = Scalar(1);
        return;
    }
    for (int i = 0; i < 9; i++) A[i] = mat[i] / fnorm;

    Scalar detB = 0, adj = 0;
    adj = A[4] * A[8] - A[5] * A[7]; detB += adj * adj;
    adj = A[3] * A[8] - A[5] * A[6]; detB += adj * adj;
    adj = A[3] * A[7] - A[4] * A[6]; detB += adj * adj;
    adj = A[1] * A[8] - A[2] * A[7]; detB += adj * adj;
    adj = A[0] * A[8] - A[2] * A[6]; detB += adj * adj;
    adj = A[0] * A[7] - A[1] * A[6]; detB += adj * adj;
    adj = A[1] * A[5] - A[2] * A[4]; detB += adj * adj;
    adj = A[0] * A[5] - A[2] * A[3]; detB += adj * adj;
    adj = A[0] * A[4] - A[1] * A[3]; detB += adj * adj;
    detB = Scalar(1) - Scalar(4) * detB;

    B[0] = -A[0] - A[4] - A[8]; B[1] = -A[5] + A[7]; B[2] = -A[6] + A[2]; B[3] = -A[1] + A[3];
    B[4] = B[1]; B[5] = -A[0] + A[4] + A[8]; B[6] = -A[1] - A[3]; B[7] = -A[2] - A[6];
    B[8] = B[2]; B[9] = B[6]; B[10] = -A[4] + A[0] + A[8]; B[11] = -A[5] - A[7];
    B[12] = B[3]; B[13] = B[7]; B[14] = B[11]; B[15] = -A[8] + A[0] + A[4];

    Scalar detA = 1, omega = 0;
    const Scalar tau = Scalar(1e-4);
    bool quick = detB <= Scalar(1) - tau;
    if (quick) {
        int rowPermute[2] = {1, 2};
        {
            int rowIndex = 0;
            Scalar largest = std::fabs(A[0]);
            if (std::fabs(A[3]) > largest) { largest = std::fabs(A[3]); rowIndex = 1; }
            if (std::fabs(A[6]) > largest) rowIndex = 2;
            if (rowIndex != 0) { detA = -detA; rowPermute[rowIndex - 1] = 0; }
            int row = rowIndex;
            Scalar divisor = A[row * 3 + 0];
            detA *= divisor;
            if (detA != Scalar(0)) {
                for (int j = 0; j < 2; j++) {
                    int subRow = rowPermute[j];
                    Scalar lower = A[subRow * 3 + 0] / divisor;
                    for (int k = 1; k < 3; k++) A[subRow * 3 + k] -= lower * A[row * 3 + k];
                }
            }
        }
        if (detA != Scalar(0)) {
            int row = 0, subRow = 0;
            if (std::fabs(A[rowPermute[0] * 3 + 1]) > std::fabs(A[rowPermute[1] * 3 + 1])) {
                row = rowPermute[0]; subRow = rowPermute[1];
            } else {
                detA = -detA; row = rowPermute[1]; subRow = rowPermute[0];
            }
            Scalar divisor = A[row * 3 + 1];
            detA *= divisor;
            Scalar lower = A[subRow * 3 + 1] / divisor;
            detA *= (A[subRow * 3 + 2] - lower * A[row * 3 + 2]);
        }
    } else {
        int rowPermute[3] = {0, 1, 2};
        int columnPermute[3] = {0, 1, 2};
        for (int i = 0; i < 2; i++) {
            Scalar largest = 0;
            int rowIndex = i, columnIndex = i;
            for (int j = i; j < 3; j++) {
                int subRow = rowPermute[j];
                for (int k = i; k < 3; k++) {
                    int subColumn = columnPermute[k];
                    Scalar entry = std::fabs(A[subRow * 3 + subColumn]);
                    if (entry > largest) { largest = entry; rowIndex = j; columnIndex = k; }
                }
            }
            if (rowPermute[i] != rowPermute[rowIndex]) { std::swap(rowPermute[i], rowPermute[rowIndex]); detA = -detA; }
            if (columnPermute[i] != columnPermute[columnIndex]) { std::swap(columnPermute[i], columnPermute[columnIndex]); detA = -detA; }
            int row = rowPermute[i], column = columnPermute[i];
            Scalar divisor = A[row * 3 + column];
            detA *= divisor;
            if (detA == Scalar(0)) break;
            for (int j = i + 1; j < 3; j++) {
                int subRow = rowPermute[j];
                Scalar lower = A[subRow * 3 + column] / divisor;
                for (int k = i + 1; k < 3; k++) {
                    int subColumn = columnPermute[k];
                    A[subRow * 3 + subColumn] -= lower * A[row * 3 + subColumn];
                }
            }
        }
        detA *= A[rowPermute[2] * 3 + columnPermute[2]];
        omega = std::fabs(A[rowPermute[1] * 3 + columnPermute[1]]);
    }

    bool notpd = detA < Scalar(0);
    if (notpd) { detA = -detA; for (int i = 0; i < 16; i++) B[i] = -B[i]; }

    const Scalar inv3 = Scalar(0.333333333333333333333333333333333333333333333333333333333333333);
    Scalar domEigen = 0;
    if (detB >= tau - inv3) {
        Scalar c = Scalar(8) * detA;
        Scalar delta0 = Scalar(1) + Scalar(3) * detB;
        Scalar delta1 = Scalar(-1) + (Scalar(27) / Scalar(16)) * c * c + Scalar(9) * detB;
        Scalar sqrtDelta0 = std::sqrt(delta0);
        Scalar alpha = delta1 / (delta0 * sqrtDelta0);
        Scalar z = (Scalar(4) * inv3) * (Scalar(1) + sqrtDelta0 * std::cos(std::acos(Clamp(alpha, Scalar(-1), Scalar(1))) * inv3));
        Scalar s = std::sqrt(z) * Scalar(0.5);
        domEigen = s + Scalar(0.5) * std::sqrt(std::max(Scalar(0), Scalar(4) - z + c / s));
    } else {
        domEigen = Scalar(1.7320508075688772935274463415058723669428052538103806280558069794519330169088);
        Scalar domEigenOld = domEigen;
        const Scalar eps = Scalar(1e-12);
        Scalar c = Scalar(8) * detA;
        do {
            Scalar p = domEigen * (domEigen * (domEigen * domEigen - Scalar(2)) - c) + detB;
            Scalar dp = domEigen * (Scalar(4) * domEigen * domEigen - Scalar(4)) - c;
            domEigenOld = domEigen;
            domEigen -= p / dp;
        } while (domEigenOld - domEigen > eps);
    }
    B[0] += domEigen; B[5] += domEigen; B[10] += domEigen; B[15] += domEigen;

    Scalar L[6];
    int indices[4] = {0, 1, 2, 3};
    {
        int diagIndex = 0, trueDiagIndex = 0;
        Scalar largest = B[0];
        for (int j = 1; j < 4; j++) {
            int ti = indices[j];
            Scalar diag = B[ti * 4 + ti];
            if (diag > largest) { largest = diag; diagIndex = j; }
        }
        if (diagIndex != 0) std::swap(indices[0], indices[diagIndex]);
        trueDiagIndex = indices[0];
        L[0] = B[trueDiagIndex * 4 + indices[1]] / largest;
        L[1] = B[trueDiagIndex * 4 + indices[2]] / largest;
        L[2] = B[trueDiagIndex * 4 + indices[3]] / largest;
        int rowIndex = 0, colIndex = 0; Scalar offDiag = 0;
        rowIndex = indices[1];
        B[rowIndex * 4 + rowIndex] -= L[0] * B[trueDiagIndex * 4 + rowIndex];
        colIndex = indices[2];
        offDiag = B[rowIndex * 4 + colIndex] - L[0] * B[trueDiagIndex * 4 + colIndex];
        B[rowIndex * 4 + colIndex] = offDiag; B[colIndex * 4 + rowIndex] = offDiag;
        colIndex = indices[3];
        offDiag = B[rowIndex * 4 + colIndex] - L[0] * B[trueDiagIndex * 4 + colIndex];
        B[rowIndex * 4 + colIndex] = offDiag; B[colIndex * 4 + rowIndex] = offDiag;
        rowIndex = indices[2];
        B[rowIndex * 4 + rowIndex] -= L[1] * B[trueDiagIndex * 4 + rowIndex];
        colIndex = indices[3];
        offDiag = B[rowIndex * 4 + colIndex] - L[1] * B[trueDiagIndex * 4 + colIndex];
        B[rowIndex * 4 + colIndex] = offDiag; B[colIndex * 4 + rowIndex] = offDiag;
        rowIndex = indices[3];
        B[rowIndex * 4 + rowIndex] -= L[2] * B[trueDiagIndex * 4 + rowIndex];
    }
    {
        int diagIndex = 1, trueDiagIndex = indices[diagIndex];
        Scalar largest = B[trueDiagIndex * 4 + trueDiagIndex];
        for (int j = 2; j < 4; j++) {
            int ti = indices[j];
            Scalar diag = B[ti * 4 + ti];
            if (diag > largest) { largest = diag; diagIndex = j; }
        }
        if (diagIndex != 1) { std::swap(indices[1], indices[diagIndex]); std::swap(L[0], L[diagIndex - 1]); }
        trueDiagIndex = indices[1];
        L[3] = B[trueDiagIndex * 4 + indices[2]] / largest;
        L[4] = B[trueDiagIndex * 4 + indices[3]] / largest;
        int rowIndex = 0, colIndex = 0; Scalar offDiag = 0;
        rowIndex = indices[2];
        B[rowIndex * 4 + rowIndex] -= L[3] * B[trueDiagIndex * 4 + rowIndex];
        colIndex = indices[3];
        offDiag = B[rowIndex * 4 + colIndex] - L[3] * B[trueDiagIndex * 4 + colIndex];
        B[rowIndex * 4 + colIndex] = offDiag; B[colIndex * 4 + rowIndex] = offDiag;
        rowIndex = indices[3];
        B[rowIndex * 4 + rowIndex] -= L[4] * B[trueDiagIndex * 4 + rowIndex];
    }

    Scalar v[4];
    if (quick) {
        if (B[indices[2] * 4 + indices[2]] < B[indices[3] * 4 + indices[3]]) {
            std::swap(indices[2], indices[3]);
            std::swap(L[1], L[2]); std::swap(L[3], L[4]);
        }
        L[5] = B[indices[2] * 4 + indices[3]] / B[indices[2] * 4 + indices[2]];
        v[indices[0]] = L[0] * L[4] + L[1] * L[5] - L[0] * L[5] * L[3] - L[2];
        v[indices[1]] = L[5] * L[3] - L[4];
        v[indices[2]] = -L[5];
        v[indices[3]] = Scalar(1);
    } else {
        Scalar dd = B[indices[2] * 4 + indices[2]] * B[indices[3] * 4 + indices[3]] -
                    B[indices[2] * 4 + indices[3]] * B[indices[2] * 4 + indices[3]];
        if (dd == Scalar(0)) {
            if (B[indices[2] * 4 + indices[2]] == Scalar(0) && B[indices[3] * 4 + indices[3]] == Scalar(0) &&
                B[indices[2] * 4 + indices[3]] == Scalar(0)) {
                v[indices[0]] = L[0] * L[4] - L[2];
                v[indices[1]] = -L[4];
                v[indices[2]] = Scalar(0); v[indices[3]] = Scalar(1);
            } else {
                Scalar aa = -B[indices[2] * 4 + indices[3]], bb = B[indices[2] * 4 + indices[2]];
                Scalar norm = std::sqrt(aa * aa + bb * bb);
                aa /= norm; bb /= norm;
                v[indices[3]] = bb; v[indices[2]] = aa;
                v[indices[1]] = -L[4] * bb - L[3] * aa;
                v[indices[0]] = -L[2] * bb - L[1] * aa - L[0] * v[indices[1]];
            }
        } else {
            Scalar ID[5];
            ID[0] = Scalar(1) / B[indices[0] * 4 + indices[0]];
            ID[1] = Scalar(1) / B[indices[1] * 4 + indices[1]];
            ID[2] = B[indices[3] * 4 + indices[3]];
            ID[3] = -B[indices[2] * 4 + indices[3]];
            ID[4] = B[indices[2] * 4 + indices[2]];
            Scalar IL[5];
            IL[0] = -L[0];
            IL[1] = L[0] * L[3] - L[1];
            IL[2] = L[0] * L[4] - L[2];
            IL[3] = -L[3];
            IL[4] = -L[4];
            const Scalar eps = Scalar(6.607e-8);
            if (omega > eps) {
                Scalar vv[4] = {IL[2], IL[4], Scalar(0), Scalar(1)};
                int nit = (int)std::ceil(Scalar(15) / (Scalar(16.86) + Scalar(2) * std::log10(omega)));
                for (int it = 0; it < nit; it++) {
                    Scalar norm = std::sqrt(vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2] + vv[3] * vv[3]);
                    for (int i = 0; i < 4; i++) vv[i] /= norm;
                    Scalar vv1 = vv[1];
                    vv[1] = IL[0] * vv[0] + vv1;
                    vv[2] = IL[1] * vv[0] + IL[3] * vv1 + vv[2];
                    vv[3] = IL[2] * vv[0] + IL[4] * vv1 + vv[3];
                    vv[0] *= ID[0]; vv[1] *= ID[1];
                    Scalar vv2 = vv[2];
                    vv[2] = (ID[2] * vv2 + ID[3] * vv[3]) / dd;
                    vv[3] = (ID[3] * vv2 + ID[4] * vv[3]) / dd;
                    vv[0] = vv[0] + IL[0] * vv[1] + IL[1] * vv[2] + IL[2] * vv[3];
                    vv[1] = vv[1] + IL[3] * vv[2] + IL[4] * vv[3];
                }
                v[indices[0]] = vv[0]; v[indices[1]] = vv[1]; v[indices[2]] = vv[2]; v[indices[3]] = vv[3];
            } else {
                Scalar vv[8];
                vv[0] = IL[1]; vv[1] = IL[3]; vv[2] = Scalar(1); vv[3] = Scalar(0);
                vv[4] = IL[2]; vv[5] = IL[4]; vv[6] = Scalar(0); vv[7] = Scalar(1);
                Orthonormalize(vv);
                for (int it = 0; it < 2; it++) {
                    Scalar vv1 = vv[1];
                    vv[1] = IL[0] * vv[0] + vv1;
                    vv[2] = IL[1] * vv[0] + IL[3] * vv1 + vv[2];
                    vv[3] = IL[2] * vv[0] + IL[4] * vv1 + vv[3];
                    Scalar vv5 = vv[5];
                    vv[5] = IL[0] * vv[4] + vv5;
                    vv[6] = IL[1] * vv[4] + IL[3] * vv5 + vv[6];
                    vv[7] = IL[2] * vv[4] + IL[4] * vv5 + vv[7];
                    vv[0] *= ID[0]; vv[1] *= ID[1];
                    Scalar vv2 = vv[2];
                    vv[2] = (ID[2] * vv2 + ID[3] * vv[3]) / dd;
                    vv[3] = (ID[3] * vv2 + ID[4] * vv[3]) / dd;
                    // numerMethod.h:928 scales vv[5] by ID[2] (reference quirk, kept)
                    vv[4] *= ID[0]; vv[5] *= ID[2];
                    Scalar vv6 = vv[6];
                    vv[6] = (ID[2] * vv6 + ID[3] * vv[7]) / dd;
                    vv[7] = (ID[3] * vv6 + ID[4] * vv[7]) / dd;
                    vv[0] = vv[0] + IL[0] * vv[1] + IL[1] * vv[2] + IL[2] * vv[3];
                    vv[1] = vv[1] + IL[3] * vv[2] + IL[4] * vv[3];
                    vv[4] = vv[4] + IL[0] * vv[5] + IL[1] * vv[6] + IL[2] * vv[7];
                    vv[5] = vv[5] + IL[3] * vv[6] + IL[4] * vv[7];
                }
                Orthonormalize(vv);
                Scalar BB[3], temp0[8], temp1[8];
                temp0[0] = vv[0] + L[0] * vv[1] + L[1] * vv[2] + L[2] * vv[3];
                temp0[1] = vv[1] + L[3] * vv[2] + L[4] * vv[3];
                temp0[2] = vv[2]; temp0[3] = vv[3];
                temp0[4] = vv[4] + L[0] * vv[5] + L[1] * vv[6] + L[2] * vv[7];
                temp0[5] = vv[5] + L[3] * vv[6] + L[4] * vv[7];
                temp0[6] = vv[6]; temp0[7] = vv[7];
                Scalar d00 = B[indices[0] * 4 + indices[0]], d11 = B[indices[1] * 4 + indices[1]];
                Scalar d22 = B[indices[2] * 4 + indices[2]], d33 = B[indices[3] * 4 + indices[3]];
                Scalar d32 = B[indices[2] * 4 + indices[3]];
                temp1[0] = temp0[0] * d00;
                temp1[1] = temp0[1] * d11;
                temp1[2] = temp0[2] * d22 + temp0[3] * d32;
                temp1[3] = temp0[2] * d32 + temp0[3] * d33;
                temp1[4] = temp0[4] * d00;
                temp1[5] = temp0[5] * d11;
                temp1[6] = temp0[6] * d22 + temp0[7] * d32;
                temp1[7] = temp0[6] * d32 + temp0[7] * d33;
                BB[0] = temp1[0] * temp0[0] + temp1[1] * temp0[1] + temp1[2] * temp0[2] + temp1[3] * temp0[3];
                BB[1] = temp1[0] * temp0[4] + temp1[1] * temp0[5] + temp1[2] * temp0[6] + temp1[3] * temp0[7];
                BB[2] = temp1[4] * temp0[4] + temp1[5] * temp0[5] + temp1[6] * temp0[6] + temp1[7] * temp0[7];
                const Scalar eps2 = Scalar(1e-15);
                if (std::fabs(BB[1]) > eps2) {
                    Scalar rr = (BB[0] - BB[2]) * Scalar(0.5);
                    Scalar w0 = rr - std::sqrt(rr * rr + BB[1] * BB[1]);
                    v[indices[0]] = vv[0] * w0 + vv[4] * BB[1];
                    v[indices[1]] = vv[1] * w0 + vv[5] * BB[1];
                    v[indices[2]] = vv[2] * w0 + vv[6] * BB[1];
                    v[indices[3]] = vv[3] * w0 + vv[7] * BB[1];
                } else {
                    if (BB[0] < BB[2]) { v[indices[0]] = vv[0]; v[indices[1]] = vv[1]; v[indices[2]] = vv[2]; v[indices[3]] = vv[3]; }
                    else { v[indices[0]] = vv[4]; v[indices[1]] = vv[5]; v[indices[2]] = vv[6]; v[indices[3]] = vv[7]; }
                }
            }
        }
    }
    Scalar norm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
    for (int i = 0; i < 4; i++) v[i] /= norm;
    Scalar v11 = v[1] * v[1], v22 = v[2] * v[2], v33 = v[3] * v[3];
    Scalar v12 = v[1] * v[2], v03 = v[0] * v[3], v13 = v[1] * v[3];
    Scalar v02 = v[0] * v[2], v23 = v[2] * v[3], v01 = v[0] * v[1];
    Scalar Q[9];
    Q[0] = Scalar(1) - Scalar(2) * (v22 + v33);
    Q[1] = Scalar(2) * (v12 - v03);
    Q[2] = Scalar(2) * (v13 + v02);
    Q[3] = Scalar(2) * (v12 + v03);
    Q[4] = Scalar(1) - Scalar(2) * (v11 + v33);
    Q[5] = Scalar(2) * (v23 - v01);
    Q[6] = Scalar(2) * (v13 - v02);
    Q[7] = Scalar(2) * (v23 + v01);
    Q[8] = Scalar(1) - Scalar(2) * (v11 + v22);
    if (notpd) for (int i = 0; i < 9; i++) Q[i] = -Q[i];
    ortho[0] = Q[0]; ortho[1] = Q[3]; ortho[2] = Q[6];
    ortho[3] = Q[1]; ortho[4] = Q[4]; ortho[5] = Q[7];
    ortho[6] = Q[2]; ortho[7] = Q[5]; ortho[8] = Q[8];
    if (sspd) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                sspd[i * 3 + j] = Q[i * 3 + 0] * mat[0 * 3 + j] + Q[i * 3 + 1] * mat[1 * 3 + j] + Q[i * 3 + 2] * mat[2 * 3 + j];
    }
}

// numerMethod.h:1019-1024 — sequential left-to-right dot
inline Scalar Dot(int width, const Scalar *x, const Scalar *y) {
    Scalar r = 0;
    for (int i = 0; i < width; i++) r += x[i] * y[i];
    return r;
}
}  // namespace oro
