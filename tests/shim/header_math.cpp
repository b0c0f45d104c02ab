// header_math.cpp -- host-side checks of the drop-in math types
// (include/gsap/quaternion.h, include/gsap/matrix3x3.h): right-handed quarter
// turns, getRotationMatrix agrees with rotateVector, normalizeQ, the product's
// composition order, Matrix3x3::invert and the matrix products.  Exit 0 when
// every check holds; no device is used.
#include <cmath>
#include <cstdio>

#include "gsap/matrix3x3.h"
#include "gsap/quaternion.h"

static int fails = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s\n", what);
        ++fails;
    }
}
static bool near(float a, float b, float tol = 1e-6f) { return std::fabs(a - b) <= tol; }
static bool near4(const float4& a, float x, float y, float z, float tol = 1e-6f) {
    return near(a.x, x, tol) && near(a.y, y, tol) && near(a.z, z, tol);
}

int main() {
    const float h = std::sqrt(0.5f);
    float4 out;
    Quaternion(0, 0, h, h).rotateVector(make_float4(1, 0, 0, 0), out);
    expect(near4(out, 0, 1, 0), "z quarter turn maps x to y");
    Quaternion(h, 0, 0, h).rotateVector(make_float4(0, 1, 0, 0), out);
    expect(near4(out, 0, 0, 1), "x quarter turn maps y to z");
    Quaternion(0, h, 0, h).rotateVector(make_float4(0, 0, 1, 0), out);
    expect(near4(out, 1, 0, 0), "y quarter turn maps z to x");

    // an arbitrary unit quaternion: matrix rows agree with rotateVector
    Quaternion q(0.1f, -0.5f, 0.3f, 0.8f);
    q.normalizeQ();
    expect(near(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w, 1.0f), "normalizeQ");
    Matrix3x3 R;
    q.getRotationMatrix(R);
    const float4 v = make_float4(0.7f, -1.3f, 2.1f, 0);
    q.rotateVector(v, out);
    const float4 Rv = R * v;
    expect(near4(Rv, out.x, out.y, out.z, 1e-5f), "R v == q v q^-1");
    // R is orthonormal: R R^T = I
    Matrix3x3 RT;
    RT.setTranspose(R);
    const Matrix3x3 I = R * RT;
    expect(near4(I[0], 1, 0, 0, 1e-6f) && near4(I[1], 0, 1, 0, 1e-6f) && near4(I[2], 0, 0, 1, 1e-6f),
           "R R^T = I");
    // (a b) rotates by b first, then a
    Quaternion a(0, 0, h, h), b(h, 0, 0, h);
    float4 t1, t2;
    b.rotateVector(make_float4(0, 1, 0, 0), t1);  // y -> z
    a.rotateVector(t1, t2);                        // z stays z
    (a * b).rotateVector(make_float4(0, 1, 0, 0), out);
    expect(near4(out, t2.x, t2.y, t2.z, 1e-6f) && near4(out, 0, 0, 1, 1e-6f), "(a b) v = a (b v a^-1) ... ");
    // inverse of a general matrix
    Matrix3x3 M(make_float4(4, 1, -2, 0), make_float4(0.5f, 3, 1, 0), make_float4(-1, 2, 5, 0));
    Matrix3x3 Mi = M;
    Mi.invert();
    const Matrix3x3 P = M * Mi;
    expect(near4(P[0], 1, 0, 0, 1e-5f) && near4(P[1], 0, 1, 0, 1e-5f) && near4(P[2], 0, 0, 1, 1e-5f),
           "M M^-1 = I");
    // row vector times matrix equals (M^T v)
    Matrix3x3 MT;
    MT.setTranspose(M);
    const float4 vm = v * M, mtv = MT * v;
    expect(near4(vm, mtv.x, mtv.y, mtv.z, 1e-5f), "v M = M^T v");
    // setAngularVelocity: rotation by |w| dt about w
    Quaternion r;
    r.setAngularVelocity(make_float4(0, 0, 3.14159265f, 0), 0.5f);  // quarter turn about z
    r.rotateVector(make_float4(1, 0, 0, 0), out);
    expect(near4(out, 0, 1, 0, 1e-6f), "setAngularVelocity quarter turn");
    static_assert(sizeof(Quaternion) == 16, "Quaternion is a float4");
    if (fails == 0) std::printf("header math: all checks passed\n");
    return fails == 0 ? 0 : 1;
}
