// gsap/matrix3x3.h -- the reference's Matrix3x3 type
// (physicboxSap2.16_Fermi/include/matrix3x3.h:36-330) for callers of the
// drop-in headers: three float4 rows (w unused), 16-byte aligned, 48 bytes,
// the layout of every inverse-inertia array of the rigid-body state.  Host
// and device.
#pragma once
#include <hip/hip_runtime.h>

class __attribute__((aligned(16))) Matrix3x3 {
public:
    __host__ __device__ Matrix3x3() {}
    __host__ __device__ Matrix3x3(const float4& r0, const float4& r1, const float4& r2) {
        m_row[0] = r0;
        m_row[1] = r1;
        m_row[2] = r2;
    }
    __host__ __device__ void setZero() {
        m_row[0] = m_row[1] = m_row[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __host__ __device__ void setIdentity() {
        setZero();
        m_row[0].x = m_row[1].y = m_row[2].z = 1.0f;
    }
    __host__ __device__ void setDiagMatrix(const float4& d) {
        setZero();
        m_row[0].x = d.x;
        m_row[1].y = d.y;
        m_row[2].z = d.z;
    }
    __host__ __device__ void setTranspose(const Matrix3x3& m) {
        const Matrix3x3 c = m;  // m may be *this
        for (int i = 0; i < 3; ++i) m_row[i] = make_float4(col(c, 0, i), col(c, 1, i), col(c, 2, i), 0.0f);
    }
    __host__ __device__ void transpose() { setTranspose(*this); }

    __host__ __device__ const float4& operator[](int i) const { return m_row[i]; }
    __host__ __device__ float4& operator[](int i) { return m_row[i]; }

    __host__ __device__ Matrix3x3& operator+=(const Matrix3x3& a) {
        for (int i = 0; i < 3; ++i) m_row[i] = add(m_row[i], a[i]);
        return *this;
    }
    __host__ __device__ Matrix3x3& operator-=(const Matrix3x3& a) {
        for (int i = 0; i < 3; ++i) m_row[i] = sub(m_row[i], a[i]);
        return *this;
    }
    __host__ __device__ Matrix3x3& operator*=(float s) {
        for (int i = 0; i < 3; ++i) m_row[i] = scale(m_row[i], s);
        return *this;
    }
    __host__ __device__ Matrix3x3& operator*=(const Matrix3x3& a) {
        *this = *this * a;
        return *this;
    }
    __host__ __device__ friend Matrix3x3 operator+(const Matrix3x3& a, const Matrix3x3& b) {
        Matrix3x3 r = a;
        return r += b;
    }
    __host__ __device__ friend Matrix3x3 operator-(const Matrix3x3& a, const Matrix3x3& b) {
        Matrix3x3 r = a;
        return r -= b;
    }
    __host__ __device__ friend Matrix3x3 operator*(const Matrix3x3& a, float s) {
        Matrix3x3 r = a;
        return r *= s;
    }
    __host__ __device__ friend Matrix3x3 operator/(const Matrix3x3& a, float s) {
        Matrix3x3 r;
        for (int i = 0; i < 3; ++i)
            r.m_row[i] = make_float4(a[i].x / s, a[i].y / s, a[i].z / s, a[i].w / s);
        return r;
    }
    // Row i of a times column j of b, summed ((x + y) + z).
    __host__ __device__ friend Matrix3x3 operator*(const Matrix3x3& a, const Matrix3x3& b) {
        Matrix3x3 r;
        for (int i = 0; i < 3; ++i)
            r.m_row[i] = make_float4(rc(a, b, i, 0), rc(a, b, i, 1), rc(a, b, i, 2), 0.0f);
        return r;
    }
    // Matrix times column vector (w of the result is 0).
    __host__ __device__ friend float4 operator*(const Matrix3x3& a, const float4& v) {
        return make_float4(dot3(a[0], v), dot3(a[1], v), dot3(a[2], v), 0.0f);
    }
    // Row vector times matrix.
    __host__ __device__ friend float4 operator*(const float4& v, const Matrix3x3& a) {
        return make_float4((v.x * a[0].x + v.y * a[1].x) + v.z * a[2].x,
                           (v.x * a[0].y + v.y * a[1].y) + v.z * a[2].y,
                           (v.x * a[0].z + v.y * a[1].z) + v.z * a[2].z, 0.0f);
    }
    // Inverse by the adjugate over the determinant (matrix3x3.h:291-312).
    __host__ __device__ void invert() {
        const Matrix3x3& m = *this;
        Matrix3x3 adj;
        adj[0] = make_float4(m[1].y * m[2].z - m[1].z * m[2].y, m[0].z * m[2].y - m[0].y * m[2].z,
                             m[0].y * m[1].z - m[0].z * m[1].y, 0.0f);
        adj[1] = make_float4(m[1].z * m[2].x - m[1].x * m[2].z, m[0].x * m[2].z - m[0].z * m[2].x,
                             m[0].z * m[1].x - m[0].x * m[1].z, 0.0f);
        adj[2] = make_float4(m[1].x * m[2].y - m[1].y * m[2].x, m[0].y * m[2].x - m[0].x * m[2].y,
                             m[0].x * m[1].y - m[0].y * m[1].x, 0.0f);
        const float det = (m[0].x * adj[0].x + m[0].y * adj[1].x) + m[0].z * adj[2].x;
        *this = adj * (1.0f / det);
    }

    float4 m_row[3];

private:
    __host__ __device__ static float dot3(const float4& a, const float4& b) {
        return (a.x * b.x + a.y * b.y) + a.z * b.z;
    }
    __host__ __device__ static float col(const Matrix3x3& m, int r, int c) {
        return c == 0 ? m[r].x : (c == 1 ? m[r].y : m[r].z);
    }
    __host__ __device__ static float rc(const Matrix3x3& a, const Matrix3x3& b, int i, int j) {
        return (a[i].x * col(b, 0, j) + a[i].y * col(b, 1, j)) + a[i].z * col(b, 2, j);
    }
    __host__ __device__ static float4 add(const float4& a, const float4& b) {
        return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    }
    __host__ __device__ static float4 sub(const float4& a, const float4& b) {
        return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
    }
    __host__ __device__ static float4 scale(const float4& a, float s) {
        return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
    }
};

static_assert(sizeof(Matrix3x3) == 48, "Matrix3x3 is three float4 rows");
