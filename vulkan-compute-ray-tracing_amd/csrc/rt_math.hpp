// rt_math.hpp -- float3 arithmetic with GLSL semantics, shared by the
// render and traversal kernels.  Every helper performs the binary32
// operations in the order the shader source writes them; the library is
// built with -ffp-contract=off so no multiply-add is fused and results are
// reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#pragma clang fp contract(off)

namespace rtamd {

struct f3 { float x, y, z; };

#define RT_HD __host__ __device__ __forceinline__

RT_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
RT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RT_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
RT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RT_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RT_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
RT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// GLSL min/max as the specification defines them.
RT_HD float gmin(float x, float y) { return y < x ? y : x; }
RT_HD float gmax(float x, float y) { return x < y ? y : x; }
RT_HD f3 gmin(f3 a, f3 b) { return f3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
RT_HD f3 gmax(f3 a, f3 b) { return f3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
RT_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
// normalize(x) = x / length(x), length = sqrt(dot(x, x)); IEEE sqrt and divide.
RT_HD f3 normalize(f3 a) { return a / sqrtf(dot(a, a)); }
RT_HD f3 reflect(f3 I, f3 N) { return I - (2.0f * dot(N, I)) * N; }
RT_HD f3 refract(f3 I, f3 N, float eta) {
    float d = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return f3{0.0f, 0.0f, 0.0f};
    return eta * I - (eta * d + sqrtf(k)) * N;
}
// Correctly rounded binary32 sine/cosine: evaluated in binary64 and rounded.
RT_HD float sin_cr(float x) { return (float)sin((double)x); }
RT_HD float cos_cr(float x) { return (float)cos((double)x); }

constexpr float kPi = 3.1415926535897932385f;   // random.glsl:1

// ------------------------------------------------------------------- RNG
// random.glsl:4-22: PCG-style step and output; one uint of state per path.
RT_HD uint32_t rng_step(uint32_t s) { return s * 747796405u + 1u; }
RT_HD float rng_next(uint32_t& state) {
    state = rng_step(state);
    uint32_t s = state;
    uint32_t word = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    word = (word >> 22) ^ word;
    return (float)word / 4294967295.0f;
}
RT_HD uint32_t rng_seed(uint32_t x, uint32_t y, uint32_t current_sample) {
    return (600u * x + y) * (current_sample + 1u);
}

RT_HD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

}  // namespace rtamd
