// FP32 arithmetic with the hardware reciprocal and square root, for the
// "mixed, fast" precision of the spherical engine (device code only).
//
// The mixed precision computes reconstruction, the Riemann fluxes and the
// EMFs in FP32.  HIP rounds FP32 division and square root correctly by
// default (a scale/reciprocal/refine/fix-up sequence of about ten
// instructions each), which the CPU oracle reproduces bit for bit.  ffloat
// is FP32 with a / b = a * rcp(b) (v_rcp_f32, within 1 ulp) and sqrt =
// v_sqrt_f32 (within 1 ulp): the same formulas (the physics templates are
// generic in their number type) at a few ulp more rounding per division.
// Its results are not the oracle's bit for bit; they are judged as the
// mixed precision is, by L1(rho), mass and energy drift and div B against
// the FP64 run.
#pragma once

#include <hip/hip_runtime.h>

#include "sph_physics.hpp"

namespace sph {

struct ffloat {
    float v;
    ffloat() = default;
    __host__ __device__ constexpr ffloat(float x) : v(x) {}
    __host__ __device__ constexpr explicit ffloat(double x) : v((float)x) {}
    __host__ __device__ constexpr explicit ffloat(int x) : v((float)x) {}
    __host__ __device__ constexpr explicit operator float() const { return v; }
    __host__ __device__ constexpr explicit operator double() const { return (double)v; }
};
static_assert(sizeof(ffloat) == sizeof(float), "ffloat is stored as float");

__device__ __forceinline__ ffloat operator+(ffloat a, ffloat b) { return ffloat(a.v + b.v); }
__device__ __forceinline__ ffloat operator-(ffloat a, ffloat b) { return ffloat(a.v - b.v); }
__device__ __forceinline__ ffloat operator*(ffloat a, ffloat b) { return ffloat(a.v * b.v); }
__device__ __forceinline__ ffloat operator/(ffloat a, ffloat b) { return ffloat(a.v * __builtin_amdgcn_rcpf(b.v)); }
__device__ __forceinline__ ffloat operator-(ffloat a) { return ffloat(-a.v); }
__device__ __forceinline__ ffloat& operator+=(ffloat& a, ffloat b) { a.v += b.v; return a; }
__device__ __forceinline__ ffloat& operator-=(ffloat& a, ffloat b) { a.v -= b.v; return a; }
__device__ __forceinline__ ffloat& operator*=(ffloat& a, ffloat b) { a.v *= b.v; return a; }
__device__ __forceinline__ bool operator<(ffloat a, ffloat b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(ffloat a, ffloat b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(ffloat a, ffloat b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(ffloat a, ffloat b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(ffloat a, ffloat b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(ffloat a, ffloat b) { return a.v != b.v; }
__device__ __forceinline__ ffloat sqrt(ffloat a) { return ffloat(__builtin_amdgcn_sqrtf(a.v)); }
__device__ __forceinline__ bool isfinite(ffloat a) { return __builtin_isfinite(a.v); }

// the solver tolerances of FP32
template <>
struct HlldTol<ffloat> {
    static constexpr float tol = HlldTol<float>::tol, ftol = HlldTol<float>::ftol, bdeg = HlldTol<float>::bdeg,
                           dp = HlldTol<float>::dp;
};

}  // namespace sph
