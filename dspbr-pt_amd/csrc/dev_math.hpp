// Device math for the gfx950 path tracer.
//
// All HIP code here is compiled with -ffp-contract=off and correctly rounded
// float division and sqrt (hipcc defaults), so these helpers give the same
// bits as the CPU oracle for +,-,*,/,sqrt. Helpers are __host__ __device__ so
// the shading code can also be compiled for the CPU (tests/devhost). Definitions that GLSL leaves to
// the implementation are pinned the same way as in the oracle:
//   normalize(v) = v * (1/sqrt(dot(v,v))), dot = (x*x'+y*y')+z*z',
//   min/max = the GLSL select forms (NaN-sensitive, unlike fminf/fmaxf),
//   mix(x,y,a) = x*(1-a) + y*a,
//   sin, cos, exp, log, pow, atan2, acos = the pinned implementations of
//   pin_math.hpp (the oracle's float build runs the same code).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pin_math.hpp"

#define DI __host__ __device__ __forceinline__
#define DD __device__ __forceinline__
// Out-of-line device functions: large leaf computations with register-only
// arguments that several BSDF lobes call (one copy in the kernel's code).
#define DN __host__ __device__ __attribute__((noinline))

// Loads through pointers that the kernels read from memory (scene tables)
// are generic (FLAT) unless the compiler is told they address global
// memory; ldg() marks them as global loads. Host builds load directly.
#ifndef DSPT_LDG
#define DSPT_LDG 1
#endif
DI float4 ldg(const float4* p) {
#if defined(__HIP_DEVICE_COMPILE__) && DSPT_LDG
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f x = *(const __attribute__((address_space(1))) v4f*)p;
  return make_float4(x.x, x.y, x.z, x.w);
#else
  return *p;
#endif
}

// Streaming (non-temporal) loads and stores of 16-byte path-state records:
// the shading pass reads each path's state once per bounce and writes the
// continuing path's next state once, so these lines are marked for early
// eviction and the L2 keeps the vertex, material and IBL lines the pass
// reuses. Host builds access memory directly.
DI float4 ld_stream(const float4* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f x = __builtin_nontemporal_load((const __attribute__((address_space(1))) v4f*)p);
  return make_float4(x.x, x.y, x.z, x.w);
#else
  return *p;
#endif
}
DI uint4 ld_stream(const uint4* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const v4u x = __builtin_nontemporal_load((const __attribute__((address_space(1))) v4u*)p);
  return make_uint4(x.x, x.y, x.z, x.w);
#else
  return *p;
#endif
}
DI void st_stream(float4* p, float4 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, (__attribute__((address_space(1))) v4f*)p);
#else
  *p = v;
#endif
}
DI void st_stream(uint4* p, uint4 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const v4u x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, (__attribute__((address_space(1))) v4u*)p);
#else
  *p = v;
#endif
}

// issued(): an empty asm that consumes the loaded values, placed after all
// loads of one record. Without it the compiler sinks loads that only one
// side of a later branch uses into that branch (e.g. a node's child boxes
// below the leaf test on its count field), which turns one memory round trip
// per record into two dependent ones. The asm loads nothing.
#ifndef DSPT_ISSUE_ALL
#define DSPT_ISSUE_ALL 1
#endif
DI void issued(const float4& a) {
#if defined(__HIP_DEVICE_COMPILE__) && DSPT_ISSUE_ALL
  asm volatile("" ::"v"(a.w));
#else
  (void)a;
#endif
}

struct f3 {
  float x, y, z;
};
DI f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
DI f3 sp3(float s) { return f3{s, s, s}; }
DI f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DI f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DI f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
DI f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
DI f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
DI f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
DI f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
DI f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }

// GLSL min/max/clamp (select forms)
DI float gmin(float x, float y) { return (y < x) ? y : x; }
DI float gmax(float x, float y) { return (x < y) ? y : x; }
DI float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
DI float sat(float x) { return gclamp(x, 0.0f, 1.0f); }
DI float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
DI f3 vmin(f3 a, f3 b) { return f3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
DI f3 vmax(f3 a, f3 b) { return f3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
DI f3 vmaxs(f3 a, float s) { return f3{gmax(a.x, s), gmax(a.y, s), gmax(a.z, s)}; }
DI f3 vclamp01(f3 v) { return vmin(vmax(v, sp3(0.0f)), sp3(1.0f)); }
DI f3 vmix(f3 x, f3 y, float a) { return f3{gmix(x.x, y.x, a), gmix(x.y, y.y, a), gmix(x.z, y.z, a)}; }
DI float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DI f3 cross3(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
DI float len3(f3 v) { return sqrtf(dot3(v, v)); }
DI f3 nrm3(f3 v) { return v * (1.0f / sqrtf(dot3(v, v))); }
DI f3 reflect3(f3 i, f3 n) { return i - (2.0f * dot3(n, i)) * n; }
DI float max3c(f3 v) { return gmax(v.x, gmax(v.y, v.z)); }
DI f3 vsqrt(f3 v) { return f3{sqrtf(v.x), sqrtf(v.y), sqrtf(v.z)}; }
DI f3 vexp(f3 v) { return f3{pm::pexp(v.x), pm::pexp(v.y), pm::pexp(v.z)}; }
DI f3 vlog(f3 v) { return f3{pm::plog(v.x), pm::plog(v.y), pm::plog(v.z)}; }
DI f3 vcos(f3 v) { return f3{pm::pcos(v.x), pm::pcos(v.y), pm::pcos(v.z)}; }
DI f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
DI f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

// SH/constants.glsl
constexpr float kPI = 3.14159265358979323f;
constexpr float kTWO_PI = 6.28318530717958648f;
constexpr float kEPS = 1e-8f;
constexpr float kEPS_COS = 0.001f;
constexpr float kEPS_PDF = 0.001f;
constexpr float kTFAR_MAX = 100000.0f;
constexpr float kBVH_FAR = 1e30f;

// Dual 16-bit multiply-with-carry generator, SH/rng.glsl:5-29.
struct Rng {
  uint32_t x, y;
  DI uint32_t next() {
    x = 36969u * (x & 65535u) + (x >> 16u);
    y = 18000u * (y & 65535u) + (y >> 16u);
    return (x << 16u) + y;
  }
  DI float uniform() { return float(next()) / 4294967296.0f; }
};
DI Rng rng_seed(int seed, int px, int py) {
  float off = float(seed * 17);
  float fx = float(px) + 0.5f, fy = float(py) + 0.5f;
  Rng r;
  r.x = (uint32_t)(397.6432f * (fx + off)) ^ (uint32_t)(32.9875f * (fy + off));
  r.y = (uint32_t)(397.6432f * (fy + 0.0f)) ^ (uint32_t)(32.9875f * (fx + 0.0f));
  return r;
}

// Wave-level (64-lane) reduction of a 64-bit counter, then one atomic.
// Must be reached by all 64 lanes of the wave.
DD void wave_add_u64(unsigned long long* dst, unsigned long long v) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    unsigned s = lo + l2;
    hi = hi + h2 + (s < lo ? 1u : 0u);
    lo = s;
  }
  unsigned long long tot = ((unsigned long long)hi << 32) | lo;
  if ((threadIdx.x & 63) == 0 && tot) atomicAdd(dst, tot);
}
