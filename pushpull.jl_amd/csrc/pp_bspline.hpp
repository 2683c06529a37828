// pp_bspline.hpp -- per-sample B-spline machinery for gfx950 kernels.
//
// A sample at source coordinate x (0-based, x = phi - 1) touches dp = degree+1 knots
// per axis starting at o = ceilf(x - dp/2) (artifacts/C/bspline.cu:66). For each axis
// this header produces, in registers:
//   w[i], dw[i], hw[i]  B-spline weight and its 1st/2nd derivative at t - i, t = x - o
//                       (bspline.cu:63-244), each already multiplied by the boundary
//                       sign of tap i (patch.cu:80,129 apply the sign to the value;
//                       a sign flip is exact, so moving it onto the weight changes no
//                       rounding),
//   off[i]              folded tap index times the axis stride (patch.cu:38-59).
//
// Fast path (|x| < 2^20, the case for any deformation of a real volume): t lies in the
// closed canonical interval [dp/2 - 1, dp/2], so every weight uses one fixed polynomial
// piece and needs no branch; indices fold with compares. Slow path (huge or extreme
// coordinates, only reachable with ext=1): the piecewise reference definitions and a
// 64-bit fold that reproduces the reference's int64 arithmetic, including x86-64's
// INT64_MIN result for a float-to-int64 conversion out of range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

constexpr int kMaxDp = 5;
constexpr float kFastLimit = 1048576.0f;  // 2^20

// ---- piecewise reference definitions (slow path and runtime-degree path) -----
// Evaluated without FMA contraction, so a sample on the large-coordinate path rounds
// every operation as the reference does.
#pragma clang fp contract(off)
__device__ __forceinline__ float bs2(float u) {
  u = fabsf(u);
  if (u < 0.5f) return 0.75f - u * u;
  if (u < 1.5f) { u = 1.5f - u; return 0.5f * u * u; }
  return 0.0f;
}
__device__ __forceinline__ float bs3(float u) {
  u = fabsf(u);
  if (u < 1.0f) return u * u * (u - 2.0f) * 0.5f + 2.0f / 3.0f;
  if (u < 2.0f) { u = 2.0f - u; return u * u * u * (1.0f / 6.0f); }
  return 0.0f;
}
__device__ __forceinline__ float bs4(float u) {
  u = fabsf(u);
  if (u < 0.5f) { u *= u; return u * (u * 0.25f - 0.625f) + 115.0f / 192.0f; }
  if (u < 1.5f)
    return u * (u * (u * (5.0f / 6.0f - u * (1.0f / 6.0f)) - 1.25f) + 5.0f / 24.0f) + 55.0f / 96.0f;
  if (u < 2.5f) { u -= 2.5f; u *= u; return u * u * (1.0f / 24.0f); }
  return 0.0f;
}
__device__ __forceinline__ float dbs2(float u) {
  float s = u > 0.0f ? 1.0f : -1.0f;
  u = fabsf(u);
  if (u < 0.5f) return -2.0f * u * s;
  if (u < 1.5f) return (u - 1.5f) * s;
  return 0.0f;
}
__device__ __forceinline__ float dbs3(float u) {
  float s = u > 0.0f ? 1.0f : -1.0f;
  u = fabsf(u);
  if (u < 1.0f) return u * (1.5f * u - 2.0f) * s;
  if (u < 2.0f) { u -= 2.0f; return -0.5f * u * u * s; }
  return 0.0f;
}
__device__ __forceinline__ float dbs4(float u) {
  float s = u > 0.0f ? 1.0f : -1.0f;
  u = fabsf(u);
  if (u < 0.5f) return (u * (u * u - 1.25f)) * s;
  if (u < 1.5f) return (u * (u * (u * (-2.0f / 3.0f) + 2.5f) - 2.5f) + 5.0f / 24.0f) * s;
  if (u < 2.5f) { u = u * 2.0f - 5.0f; return (1.0f / 48.0f) * u * u * u * s; }
  return 0.0f;
}
__device__ __forceinline__ float hbs3(float u) {
  u = fabsf(u);
  if (u < 1.0f) return 3.0f * u - 2.0f;
  if (u < 2.0f) return 2.0f - u;
  return 0.0f;
}
__device__ __forceinline__ float hbs4(float u) {
  u = fabsf(u);
  if (u < 0.5f) return 3.0f * u * u - 1.25f;
  if (u < 1.5f) return u * (-2.0f * u + 5.0f) - 2.5f;
  if (u < 2.5f) { u = u * 2.0f - 5.0f; return u * u * 0.125f; }
  return 0.0f;
}

// Weights for runtime dp (1..5) using the piecewise forms.
__device__ __forceinline__ void weights_rt(int dp, float t, float* w, float* dw, float* hw) {
#pragma unroll
  for (int k = 0; k < kMaxDp; ++k) {
    const float u = t - (float)k;
    float a = 0.f, b = 0.f, c = 0.f;
    switch (dp) {
      case 1: a = (k == 0) ? 1.f : 0.f; break;
      case 2: a = (k == 0) ? 1.f - t : (k == 1 ? t : 0.f); b = (k == 0) ? -1.f : (k == 1 ? 1.f : 0.f); break;
      case 3: a = bs2(u); b = dbs2(u); c = (k == 0 || k == 2) ? 1.f : (k == 1 ? -2.f : 0.f); break;
      case 4: a = bs3(u); b = dbs3(u); c = hbs3(u); break;
      default: a = bs4(u); b = dbs4(u); c = hbs4(u); break;
    }
    w[k] = a; dw[k] = b; hw[k] = c;
  }
}

#pragma clang fp contract(fast)

// Weights for compile-time DP with t in the canonical interval [DP/2-1, DP/2]: one
// polynomial piece per tap (the same pieces the piecewise forms select there).
template <int DP>
__device__ __forceinline__ void weights_ct(float t, float* w, float* dw, float* hw) {
  if constexpr (DP == 1) {
    w[0] = 1.f; dw[0] = 0.f; hw[0] = 0.f;
  } else if constexpr (DP == 2) {
    w[0] = 1.f - t; w[1] = t;
    dw[0] = -1.f; dw[1] = 1.f;
    hw[0] = 0.f; hw[1] = 0.f;
  } else if constexpr (DP == 3) {        // t in [0.5, 1.5]
    const float a = 1.5f - t, c = t - 0.5f, b = t - 1.0f;
    w[0] = 0.5f * a * a;  w[1] = 0.75f - b * b;  w[2] = 0.5f * c * c;
    dw[0] = -a;           dw[1] = -2.f * b;      dw[2] = c;
    hw[0] = 1.f;          hw[1] = -2.f;          hw[2] = 1.f;
  } else if constexpr (DP == 4) {        // t in [1, 2]
    const float a = 2.f - t, b = t - 1.f;  // a + b = 1
    w[0] = a * a * a * (1.f / 6.f);
    w[1] = b * b * (b - 2.f) * 0.5f + 2.f / 3.f;
    w[2] = a * a * (a - 2.f) * 0.5f + 2.f / 3.f;
    w[3] = b * b * b * (1.f / 6.f);
    dw[0] = -0.5f * a * a;
    dw[1] = b * (1.5f * b - 2.f);
    dw[2] = -(a * (1.5f * a - 2.f));
    dw[3] = 0.5f * b * b;
    hw[0] = a; hw[1] = 3.f * b - 2.f; hw[2] = 3.f * a - 2.f; hw[3] = b;
  } else {                                // DP == 5, t in [1.5, 2.5]
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const float u = t - (float)k;
      w[k] = bs4(u); dw[k] = dbs4(u); hw[k] = hbs4(u);
    }
  }
}

// x86-64 float->int64 conversion of the reference (cvttss2si): INT64_MIN when out of range.
__device__ __forceinline__ long long knot64(float c) {
  if (!(c >= -9223372036854775808.0f && c < 9223372036854775808.0f)) return (long long)0x8000000000000000ULL;
  return (long long)c;
}

// The divisor of a modulo on the multi-period path, made opaque to the optimiser (an empty
// asm statement on it), so that the division's setup -- a reciprocal of the divisor --
// stays on that rarely taken path instead of being hoisted into every kernel's or tile's
// prologue, where it would run on every call.
__device__ __forceinline__ int rare_divisor(int d) {
  asm volatile("" : "+v"(d));
  return d;
}

// Fold index i onto [0, d): returns the folded index, sets neg for a mirrored copy under
// code 2 (patch.cu:5-59; any code other than 0 and 2 reflects like code 1).
__device__ __forceinline__ int fold32(int code, int i, int d, bool& neg) {
  neg = false;
  if ((unsigned)i < (unsigned)d) return i;
  if (code == 0) {
    if (i < 0 && i >= -d) return i + d;
    if (i >= d && i < 2 * d) return i - d;
    const int dd = rare_divisor(d);
    int r = i % dd;
    return r < 0 ? r + dd : r;
  }
  int m;
  if (i < 0 && i >= -d) m = i + 2 * d;
  else if (i >= d && i < 2 * d) m = i;
  else {
    const int d2 = rare_divisor(2 * d);
    m = i % d2;
    if (m < 0) m += d2;
  }
  if (m < d) return m;
  neg = (code == 2);
  return 2 * d - 1 - m;
}

// 64-bit fold for the large-coordinate path; returns the folded index in the low 32
// bits and the antisymmetric sign flag in bit 32.
__device__ __noinline__ long long fold64(int code, long long i, long long d) {
  if (code == 0) {
    long long r = i % d;
    if (r < 0) r += d;
    return r;
  }
  const long long p = 2 * d;
  long long m = i % p;
  if (m < 0) m += p;
  if (m < d) return m;
  return (p - 1 - m) | ((code == 2) ? (1LL << 32) : 0LL);
}

// Weight sets an op needs
enum : int { kW = 1, kDW = 2, kHW = 4 };

// One axis of a sample: DP > 0 compile-time taps; DP == 0 runtime dp (<= 5), taps
// beyond n are never touched by the callers.
template <int DP>
struct Axis {
  static constexpr int kDP = DP;
  static constexpr int NP = DP > 0 ? DP : kMaxDp;
  float w[NP], dw[NP], hw[NP];
  int off[NP];
  int n;
  bool fast;  // |x| < 2^20: canonical weights, 32-bit fold
};

template <int DP, int NEED>
__device__ __forceinline__ void make_axis(float x, int dp_rt, int d, int stride, int code, Axis<DP>& ax) {
  constexpr int NP = Axis<DP>::NP;
  const int dp = DP > 0 ? DP : dp_rt;
  ax.n = dp;
  const float c = ceilf(x - 0.5f * (float)dp);
  float w[NP], dw[NP], hw[NP];
  const bool fast = fabsf(x) < kFastLimit;
  ax.fast = fast;
  int o32 = 0;
  long long o64 = 0;
  float t;
  if (fast) {
    o32 = (int)c;
    t = x - (float)o32;
    if constexpr (DP > 0) weights_ct<DP>(t, w, dw, hw);
    else weights_rt(dp, t, w, dw, hw);
  } else {
    o64 = knot64(c);
    t = x - (float)o64;
    float w5[kMaxDp], dw5[kMaxDp], hw5[kMaxDp];
    weights_rt(dp, t, w5, dw5, hw5);
#pragma unroll
    for (int i = 0; i < NP; ++i) { w[i] = w5[i]; dw[i] = dw5[i]; hw[i] = hw5[i]; }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    bool neg = false;
    int b = 0;
    if (DP > 0 || i < dp) {
      if (fast) {
        b = fold32(code, o32 + i, d, neg);
      } else {
        const long long r = fold64(code, o64 + i, d);
        b = (int)(r & 0xffffffffLL);
        neg = (r >> 32) != 0;
      }
    }
    const float s = neg ? -1.0f : 1.0f;
    ax.off[i] = b * stride;
    if (NEED & kW) ax.w[i] = w[i] * s;
    if (NEED & kDW) ax.dw[i] = dw[i] * s;
    if (NEED & kHW) ax.hw[i] = hw[i] * s;
  }
}

// VOXOK (artifacts/C/pushpull.c:6-9), in FP32 exactly as the reference.
__device__ __forceinline__ bool voxok(float x, float y, float z, int ext, float lx, float ly, float lz) {
  const bool fin = isfinite(x) && isfinite(y) && isfinite(z);
  if (!fin) return false;
  if (ext) return true;
  return x >= -0.01f && x <= lx && y >= -0.01f && y <= ly && z >= -0.01f && z <= lz;
}

}  // namespace pp
