// Device-side 2-D vector arithmetic with the rounding semantics of the UMANS reference
// (src/Engine/tools/vector2D.h). The library is compiled with -ffp-contract=off, so
// every float expression below rounds exactly as the reference's scalar SSE code does;
// values the reference computes in double are computed in double here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define UD __device__ __forceinline__

namespace umans {

constexpr float kMaxFloat = 3.402823466e+38f;          // costFunction.h:43 (numeric_limits<float>::max)
constexpr double kPI = 3.1415926535897;                // vector2D.h:28

struct V2 {
	float x, y;
};

// std::min / std::max semantics (first argument wins ties; NaN handling as in libstdc++)
UD float smin(float a, float b) { return (b < a) ? b : a; }
UD float smax(float a, float b) { return (a < b) ? b : a; }

UD V2 v2(float x, float y) { return V2{x, y}; }
UD V2 v2(float2 a) { return V2{a.x, a.y}; }
UD float2 f2(V2 a) { return make_float2(a.x, a.y); }
UD V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
UD V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
UD V2 operator-(V2 a) { return V2{-a.x, -a.y}; }
UD V2 operator*(float s, V2 v) { return V2{v.x * s, v.y * s}; }   // vector2D.h:122-125
UD V2 operator*(V2 v, float s) { return V2{v.x * s, v.y * s}; }
UD V2 operator/(V2 v, float s) { const float r = 1.0f / s; return V2{v.x * r, v.y * r}; }  // vector2D.h:132-135
UD float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
UD float sqr_mag(V2 a) { return a.x * a.x + a.y * a.y; }
UD float mag(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
UD bool is_zero(V2 a) { return a.x == 0.0f && a.y == 0.0f; }
// vector2D.h:55-72: true division by the magnitude
UD V2 normalized(V2 a) {
	const float m = mag(a);
	if (m > 0.0f) return V2{a.x / m, a.y / m};
	return a;
}
UD float det(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }

// vector2D.h:166-170 with float cos/sin (the reference binds the float overloads),
// double products, cast back to float.
UD V2 rotate_cs(V2 v, float c, float s) {
	const double C = (double)c, S = (double)s;
	return V2{(float)(C * (double)v.x - S * (double)v.y), (float)(S * (double)v.x + C * (double)v.y)};
}
UD V2 rotate_ccw(V2 v, float radians) { return rotate_cs(v, cosf(radians), sinf(radians)); }

// vector2D.h:177-190: float magnitudes and dot, double quotient and acos
UD float angle_between(V2 a, V2 b) {
	const double lengths = (double)(mag(a) * mag(b));
	if (lengths == 0.0) return 0.0f;
	const double frac = (double)dot(a, b) / lengths;
	if (frac < -1.0 || frac > 1.0) return 0.0f;
	return (float)acos(frac);
}
// vector2D.h:192-199
UD float cos_angle(V2 a, V2 b) {
	const double lengths = (double)(mag(a) * mag(b));
	if (lengths == 0.0) return 0.0f;
	return (float)((double)dot(a, b) / lengths);
}
// vector2D.h:201-219
UD float ccw_angle(V2 a, V2 b) {
	const float ang = angle_between(a, b);
	const float cross = a.x * b.y - a.y * b.x;
	return cross < 0.0f ? -ang : ang;
}
// vector2D.h:241-247
UD V2 clamp_vec(V2 v, float max_len) {
	if (mag(v) <= max_len) return v;
	return normalized(v) * max_len;
}
// vector2D.h:249-272 (all double)
UD bool line_intersection(V2 a, V2 b, V2 c, V2 d, V2& res) {
	const double a1 = (double)(b.y - a.y), b1 = (double)(a.x - b.x);
	const double c1 = a1 * (double)a.x + b1 * (double)a.y;
	const double a2 = (double)(d.y - c.y), b2 = (double)(c.x - d.x);
	const double c2 = a2 * (double)c.x + b2 * (double)c.y;
	const double dt = a1 * b2 - a2 * b1;
	if (fabs(dt) < 0.00001) return false;
	res = V2{(float)((b2 * c1 - b1 * c2) / dt), (float)((a1 * c2 - a2 * c1) / dt)};
	return true;
}

// std::minstd_rand0 (multiplier 16807, modulus 2^31-1) feeding
// std::uniform_real_distribution<float> as libstdc++ implements it: one draw per
// value, u = float(x - 1) / float(2^31 - 2) (the divisor rounds to 2^31), clamped
// below 1, then u * (hi - lo) + lo in float. Agent::ComputeRandomNumber, agent.cpp:131-135.
UD float rng_uniform(uint32_t& state, float lo, float hi) {
	state = (uint32_t)(((uint64_t)state * 16807ull) % 2147483647ull);
	float u = (float)(state - 1u) / 2147483648.0f;
	if (u >= 1.0f) u = 0.99999994f;   // nextafter(1, 0)
	return u * (hi - lo) + lo;
}

} // namespace umans
