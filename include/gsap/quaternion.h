// gsap/quaternion.h -- the reference's Quaternion type
// (physicboxSap2.16_Fermi/include/quaternion.h:25-148) for callers of the
// drop-in headers: a float4 (x, y, z, w) with w the scalar part, so arrays of
// it are the float4 arrays the rigid-body state holds.  Host and device.
//
// Conventions kept from the reference: the Hamilton product written as
// cross(a, b) + a.w b + b.w a with scalar a.w b.w - dot3(a, b); a vector v is
// rotated as q v q^-1 with q^-1 the conjugate (q is kept unit length); the
// rotation matrix rows are those of the same rotation.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "matrix3x3.h"

struct Quaternion : public float4 {
    __host__ __device__ Quaternion() : float4(make_float4(0.0f, 0.0f, 0.0f, 1.0f)) {}
    __host__ __device__ Quaternion(const float4& a) : float4(a) {}
    __host__ __device__ Quaternion(float x_, float y_, float z_, float w_)
        : float4(make_float4(x_, y_, z_, w_)) {}

    // Rotation by |w| dt about w (quaternion.h:59-76); identity below 1e-6 rad.
    __host__ __device__ void setAngularVelocity(const float4& w, float dt) {
        const float ax = w.x * dt, ay = w.y * dt, az = w.z * dt;
        const float ang = sqrtf((ax * ax + ay * ay) + az * az);
        if (ang < 1e-6f) {
            x = y = z = 0.0f;
            this->w = 1.0f;
            return;
        }
        const float s = sinf(0.5f * ang) / ang;
        x = ax * s;
        y = ay * s;
        z = az * s;
        this->w = cosf(0.5f * ang);
    }
    // Vector part of a x b (quaternion.h:78-84).
    __host__ __device__ void setCross3(const Quaternion& a, const Quaternion& b) {
        x = a.y * b.z - a.z * b.y;
        y = a.z * b.x - a.x * b.z;
        z = a.x * b.y - a.y * b.x;
        w = 0.0f;
    }
    // Rotation matrix of a unit quaternion (quaternion.h:86-106).
    __host__ __device__ void getRotationMatrix(Matrix3x3& out) const {
        const float xx = x * x, yy = y * y, zz = z * z;
        out[0] = make_float4(1 - 2 * yy - 2 * zz, 2 * x * y - 2 * w * z, 2 * x * z + 2 * w * y, 0.0f);
        out[1] = make_float4(2 * x * y + 2 * w * z, 1 - 2 * xx - 2 * zz, 2 * y * z - 2 * w * x, 0.0f);
        out[2] = make_float4(2 * x * z - 2 * w * y, 2 * y * z + 2 * w * x, 1 - 2 * xx - 2 * yy, 0.0f);
    }
    // Conjugate (the inverse of a unit quaternion, quaternion.h:108-114).
    __host__ __device__ void invert() {
        x = -x;
        y = -y;
        z = -z;
    }
    // Scale to unit length (quaternion.h:116-128).
    __host__ __device__ void normalizeQ() {
        const float inv = 1.0f / sqrtf(((x * x + y * y) + z * z) + w * w);
        x *= inv;
        y *= inv;
        z *= inv;
        w *= inv;
    }
    __host__ __device__ friend Quaternion operator*(const Quaternion& a, const Quaternion& b) {
        Quaternion r;
        r.x = (a.y * b.z - a.z * b.y) + (a.w * b.x + b.w * a.x);
        r.y = (a.z * b.x - a.x * b.z) + (a.w * b.y + b.w * a.y);
        r.z = (a.x * b.y - a.y * b.x) + (a.w * b.z + b.w * a.z);
        r.w = a.w * b.w - ((a.x * b.x + a.y * b.y) + a.z * b.z);
        return r;
    }
    // out = q in q^-1 (quaternion.h:141-148); in.w is carried through the
    // products the way the reference's is.
    __host__ __device__ void rotateVector(const float4& in, float4& out) const {
        Quaternion inv = *this;
        inv.invert();
        out = (*this * Quaternion(in)) * inv;
    }
};
