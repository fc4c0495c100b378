// Small vector / affine-matrix algebra shared by host and device code.
//
// Semantics follow the reference's math (src/math/Mat3x4f.hpp:25-267, src/math/Vector.hpp:139-196):
// row-major 3x4 affine matrices, rotationRad(x,y,z) composed in the DirectXMath order, TRS = T*R(deg)*S,
// and a 4x4 cofactor inverse that returns all zeros for a singular matrix. The formulas are the standard
// closed forms; they are restated here so that results match the reference bit-for-bit where the reference
// tests compare exactly (axis-aligned transforms).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RGL_HD __host__ __device__ __forceinline__
#else
#define RGL_HD inline
#endif

namespace rgl {

struct Vec2f
{
	float x, y;
};

struct Vec3i
{
	int32_t x, y, z;
};

struct Vec3f
{
	float x, y, z;
	RGL_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
	RGL_HD float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

RGL_HD Vec3f make3(float x, float y, float z) { return Vec3f{x, y, z}; }
RGL_HD Vec3f operator+(Vec3f a, Vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RGL_HD Vec3f operator-(Vec3f a, Vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RGL_HD Vec3f operator*(Vec3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RGL_HD Vec3f operator*(float s, Vec3f a) { return {a.x * s, a.y * s, a.z * s}; }
RGL_HD Vec3f operator/(Vec3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RGL_HD float dot(Vec3f a, Vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RGL_HD Vec3f cross(Vec3f a, Vec3f b)
{
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RGL_HD float length(Vec3f a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// Component-wise division by the length, as src/math/Vector.hpp:152 does.
RGL_HD Vec3f normalized(Vec3f a)
{
	float l = length(a);
	return {a.x / l, a.y / l, a.z / l};
}

struct Mat3x4f
{
	float rc[3][4];

	static RGL_HD Mat3x4f identity() { return {{{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}}}; }
	static RGL_HD Mat3x4f zero() { return {{{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}}}; }
	static RGL_HD Mat3x4f translation(float x, float y, float z) { return {{{1, 0, 0, x}, {0, 1, 0, y}, {0, 0, 1, z}}}; }
	static RGL_HD Mat3x4f scale(float x, float y, float z) { return {{{x, 0, 0, 0}, {0, y, 0, 0}, {0, 0, z, 0}}}; }

	// Euler composition Rz(z)*Ry(y)*Rx(x) written out (src/math/Mat3x4f.hpp:38-60).
	static RGL_HD Mat3x4f rotationRad(float x, float y, float z)
	{
		float cr = cosf(x), sr = sinf(x);
		float cp = cosf(y), sp = sinf(y);
		float cy = cosf(z), sy = sinf(z);
		Mat3x4f m = identity();
		m.rc[0][0] = cp * cy;
		m.rc[0][1] = sr * sp * cy - cr * sy;
		m.rc[0][2] = sr * sy + cr * sp * cy;
		m.rc[1][0] = cp * sy;
		m.rc[1][1] = cr * cy + sr * sp * sy;
		m.rc[1][2] = cr * sp * sy - sr * cy;
		m.rc[2][0] = -sp;
		m.rc[2][1] = sr * cp;
		m.rc[2][2] = cr * cp;
		return m;
	}

	// Rotation about one axis (src/math/Mat3x4f.hpp:64-70); axis uses rgl_axis_t values 1..3.
	static RGL_HD Mat3x4f rotationAxisRad(int axis, float angle)
	{
		return rotationRad(axis == 1 ? angle : 0.0f, axis == 2 ? angle : 0.0f, axis == 3 ? angle : 0.0f);
	}

	static RGL_HD Mat3x4f rotationDeg(float x, float y, float z)
	{
		const float k = static_cast<float>(M_PI) / 180.0f;
		return rotationRad(x * k, y * k, z * k);
	}

	RGL_HD Vec3f translationVec() const { return {rc[0][3], rc[1][3], rc[2][3]}; }

	RGL_HD Mat3x4f rotation() const
	{
		return {{{rc[0][0], rc[0][1], rc[0][2], 0}, {rc[1][0], rc[1][1], rc[1][2], 0}, {rc[2][0], rc[2][1], rc[2][2], 0}}};
	}

	// Euler extraction for the Unity-style ZXY left-handed convention (src/math/Mat3x4f.hpp:146-175).
	RGL_HD float toRotationXOrderZXYLeftHandRad() const
	{
		if (rc[1][2] < 1) {
			if (rc[1][2] > -1) {
				return asinf(rc[1][2]);
			}
			return -static_cast<float>(M_PI) / 2.0f;
		}
		return static_cast<float>(M_PI) / 2.0f;
	}
	RGL_HD float toRotationYOrderZXYLeftHandRad() const
	{
		if (rc[2][1] < 1 && rc[2][1] > -1) {
			return atan2f(-rc[2][0], rc[2][2]);
		}
		return 0.0f;
	}

	// Affine inverse through the 4x4 cofactor expansion; zero matrix if singular (src/math/Mat3x4f.hpp:188-245).
	RGL_HD Mat3x4f inverse() const
	{
		const float m00 = rc[0][0], m01 = rc[0][1], m02 = rc[0][2], m03 = rc[0][3];
		const float m10 = rc[1][0], m11 = rc[1][1], m12 = rc[1][2], m13 = rc[1][3];
		const float m20 = rc[2][0], m21 = rc[2][1], m22 = rc[2][2], m23 = rc[2][3];
		// With the implicit last row (0,0,0,1) the 2x2 minors that involve row 3 collapse to single entries
		// of row 2 (or to zero), so the 4x4 cofactor expansion reduces to the 3x3 adjugate below.
		const float det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
		if (det == 0.0f) {
			return zero();
		}
		const float id = 1.0f / det;
		Mat3x4f r;
		r.rc[0][0] = id * (m11 * m22 - m12 * m21);
		r.rc[0][1] = id * -(m01 * m22 - m02 * m21);
		r.rc[0][2] = id * (m01 * m12 - m02 * m11);
		r.rc[1][0] = id * -(m10 * m22 - m12 * m20);
		r.rc[1][1] = id * (m00 * m22 - m02 * m20);
		r.rc[1][2] = id * -(m00 * m12 - m02 * m10);
		r.rc[2][0] = id * (m10 * m21 - m11 * m20);
		r.rc[2][1] = id * -(m00 * m21 - m01 * m20);
		r.rc[2][2] = id * (m00 * m11 - m01 * m10);
		// Translation column: -(R^-1 * t) expanded through the same cofactors.
		const float A2312 = m12 * m23 - m13 * m22, A1312 = m11 * m23 - m13 * m21, A1212 = m11 * m22 - m12 * m21;
		const float A0312 = m10 * m23 - m13 * m20, A0212 = m10 * m22 - m12 * m20, A0112 = m10 * m21 - m11 * m20;
		r.rc[0][3] = id * -(m01 * A2312 - m02 * A1312 + m03 * A1212);
		r.rc[1][3] = id * (m00 * A2312 - m02 * A0312 + m03 * A0212);
		r.rc[2][3] = id * -(m00 * A1312 - m01 * A0312 + m03 * A0112);
		return r;
	}
};

// Three-term dot product a0*b0 + a1*b1 + a2*b2. In device code every rounding is explicit (two fused multiply-adds
// in a fixed order), so a matrix product gives the same bits in every kernel whatever the compiler would contract
// around it: batched and single launches of the same ray produce identical poses, angles and hits.
RGL_HD float dot3(float a0, float b0, float a1, float b1, float a2, float b2)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return __fmaf_rn(a2, b2, __fmaf_rn(a1, b1, __fmul_rn(a0, b0)));
#else
	return (a0 * b0) + (a1 * b1) + (a2 * b2);
#endif
}

RGL_HD float add1(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return __fadd_rn(a, b);
#else
	return a + b;
#endif
}

RGL_HD Mat3x4f operator*(const Mat3x4f& a, const Mat3x4f& b)
{
	Mat3x4f r;
	for (int y = 0; y < 3; ++y) {
		for (int x = 0; x < 4; ++x) {
			r.rc[y][x] = dot3(a.rc[y][0], b.rc[0][x], a.rc[y][1], b.rc[1][x], a.rc[y][2], b.rc[2][x]);
		}
		r.rc[y][3] = add1(r.rc[y][3], a.rc[y][3]);
	}
	return r;
}

RGL_HD Vec3f operator*(const Mat3x4f& m, Vec3f v)
{
	return {add1(dot3(m.rc[0][0], v.x, m.rc[0][1], v.y, m.rc[0][2], v.z), m.rc[0][3]),
	        add1(dot3(m.rc[1][0], v.x, m.rc[1][1], v.y, m.rc[1][2], v.z), m.rc[1][3]),
	        add1(dot3(m.rc[2][0], v.x, m.rc[2][1], v.y, m.rc[2][2], v.z), m.rc[2][3])};
}

// Linear part only (direction transform).
RGL_HD Vec3f mulDir(const Mat3x4f& m, Vec3f v)
{
	return {dot3(m.rc[0][0], v.x, m.rc[0][1], v.y, m.rc[0][2], v.z),
	        dot3(m.rc[1][0], v.x, m.rc[1][1], v.y, m.rc[1][2], v.z),
	        dot3(m.rc[2][0], v.x, m.rc[2][1], v.y, m.rc[2][2], v.z)};
}

// TRS = T * R(deg) * S (src/math/Mat3x4f.hpp:91-97).
RGL_HD Mat3x4f TRS(Vec3f t, Vec3f rDeg, Vec3f s)
{
	return Mat3x4f::translation(t.x, t.y, t.z) * Mat3x4f::rotationDeg(rDeg.x, rDeg.y, rDeg.z) *
	       Mat3x4f::scale(s.x, s.y, s.z);
}

// ------------------------------------------------------------------------------------------------ FP64 composition
// Sensor poses and entity transforms at large world coordinates lose their small offsets when composed in FP32
// (a ray at 40 m from the origin has its direction rounded to ~4e-6 if it is formed as M(0,0,1) - M(0,0,0)). Poses
// are therefore composed in FP64 and rounded to FP32 once; the FP64 origin is kept for the world -> object transform
// and for the hit distance. Device code fixes every rounding (fused multiply-adds in a fixed order), so a ray gets the
// same bits in every kernel.
struct Vec3d
{
	double x, y, z;
};

struct Mat3x4d
{
	double rc[3][4];
};

RGL_HD double fma64(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return __fma_rn(a, b, c);
#else
	return std::fma(a, b, c);
#endif
}

// a0*b0 + a1*b1 + a2*b2 + c with the same rounding order everywhere.
RGL_HD double dot3d(double a0, double b0, double a1, double b1, double a2, double b2, double c)
{
	return fma64(a2, b2, fma64(a1, b1, fma64(a0, b0, c)));
}

RGL_HD Mat3x4d widen(const Mat3x4f& m)
{
	Mat3x4d r;
	for (int y = 0; y < 3; ++y) {
		for (int x = 0; x < 4; ++x) {
			r.rc[y][x] = m.rc[y][x];
		}
	}
	return r;
}

RGL_HD Mat3x4f round32(const Mat3x4d& m)
{
	Mat3x4f r;
	for (int y = 0; y < 3; ++y) {
		for (int x = 0; x < 4; ++x) {
			r.rc[y][x] = static_cast<float>(m.rc[y][x]);
		}
	}
	return r;
}

RGL_HD Mat3x4d operator*(const Mat3x4d& a, const Mat3x4d& b)
{
	Mat3x4d r;
	for (int y = 0; y < 3; ++y) {
		for (int x = 0; x < 4; ++x) {
			r.rc[y][x] = dot3d(a.rc[y][0], b.rc[0][x], a.rc[y][1], b.rc[1][x], a.rc[y][2], b.rc[2][x],
			                   x == 3 ? a.rc[y][3] : 0.0);
		}
	}
	return r;
}

RGL_HD Vec3d apply(const Mat3x4d& m, Vec3d p)
{
	return {dot3d(m.rc[0][0], p.x, m.rc[0][1], p.y, m.rc[0][2], p.z, m.rc[0][3]),
	        dot3d(m.rc[1][0], p.x, m.rc[1][1], p.y, m.rc[1][2], p.z, m.rc[1][3]),
	        dot3d(m.rc[2][0], p.x, m.rc[2][1], p.y, m.rc[2][2], p.z, m.rc[2][3])};
}

RGL_HD Vec3d apply(const Mat3x4f& m, Vec3d p) { return apply(widen(m), p); }

RGL_HD Vec3f round32(Vec3d v) { return {static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z)}; }

// Affine inverse in FP64 (adjugate over the determinant); zero matrix if singular.
RGL_HD Mat3x4d inverse64(const Mat3x4d& a)
{
	const double c00 = a.rc[1][1] * a.rc[2][2] - a.rc[1][2] * a.rc[2][1];
	const double c01 = a.rc[1][2] * a.rc[2][0] - a.rc[1][0] * a.rc[2][2];
	const double c02 = a.rc[1][0] * a.rc[2][1] - a.rc[1][1] * a.rc[2][0];
	const double det = a.rc[0][0] * c00 + a.rc[0][1] * c01 + a.rc[0][2] * c02;
	Mat3x4d r{};
	if (det == 0.0) {
		return r;
	}
	const double id = 1.0 / det;
	r.rc[0][0] = c00 * id;
	r.rc[0][1] = (a.rc[0][2] * a.rc[2][1] - a.rc[0][1] * a.rc[2][2]) * id;
	r.rc[0][2] = (a.rc[0][1] * a.rc[1][2] - a.rc[0][2] * a.rc[1][1]) * id;
	r.rc[1][0] = c01 * id;
	r.rc[1][1] = (a.rc[0][0] * a.rc[2][2] - a.rc[0][2] * a.rc[2][0]) * id;
	r.rc[1][2] = (a.rc[0][2] * a.rc[1][0] - a.rc[0][0] * a.rc[1][2]) * id;
	r.rc[2][0] = c02 * id;
	r.rc[2][1] = (a.rc[0][1] * a.rc[2][0] - a.rc[0][0] * a.rc[2][1]) * id;
	r.rc[2][2] = (a.rc[0][0] * a.rc[1][1] - a.rc[0][1] * a.rc[1][0]) * id;
	for (int y = 0; y < 3; ++y) {
		r.rc[y][3] = -(r.rc[y][0] * a.rc[0][3] + r.rc[y][1] * a.rc[1][3] + r.rc[y][2] * a.rc[2][3]);
	}
	return r;
}

static_assert(sizeof(Mat3x4f) == 48, "Mat3x4f must match rgl_mat3x4f");
static_assert(sizeof(Vec3f) == 12, "Vec3f must match rgl_vec3f");

} // namespace rgl
