// pp_sample.hpp -- per-sample gather / scatter building blocks shared by the
// one-work-item-per-voxel kernels (pp_kernels.hip) and the tiled kernels (pp_tiled.hip).
#pragma once
#include "pp_bspline.hpp"
#include "pp_kernels.hpp"

namespace pp {


// Source coordinate of output voxel (i,j,k) under a 0-based affine, in the reference
// CPU order xj = (A3 j + A6 k) + A9, x = xj + A0 i (artifacts/C/pushpull.c:190-200),
// rounded per operation so knots and masks match it exactly.
#pragma clang fp contract(off)
__device__ __forceinline__ void affine_x(const float* A, int i, int j, int k, float& x, float& y, float& z) {
  const float fi = (float)i, fj = (float)j, fk = (float)k;
  x = (A[3] * fj + A[6] * fk + A[9]) + A[0] * fi;
  y = (A[4] * fj + A[7] * fk + A[10]) + A[1] * fi;
  z = (A[5] * fj + A[8] * fk + A[11]) + A[2] * fi;
}
#pragma clang fp contract(fast)

// Coordinates (0-based) of sample v at grid index (i, j, k): phi - 1 (pushpull.c:24-26),
// the affine map, or, in displacement mode, (id + dscale * u) - 1 with id the 1-based
// identity (src/compositions.jl:45: Id .+ u, then pull's phi - 1).
// ... from the deformation's three values p0, p1, p2 at the sample (not the affine map).
__device__ __forceinline__ void coords_from(const KArgs& a, float p0, float p1, float p2, int i, int j, int k,
                                            float& x, float& y, float& z) {
  if (a.disp) {
    x = ((float)(i + 1) + a.dscale * p0) - 1.0f;
    y = ((float)(j + 1) + a.dscale * p1) - 1.0f;
    z = ((float)(k + 1) + a.dscale * p2) - 1.0f;
    return;
  }
  x = p0 - 1.0f;
  y = p1 - 1.0f;
  z = p2 - 1.0f;
}

__device__ __forceinline__ void sample_coords_at(const KArgs& a, const float* phi, int v, int i, int j, int k,
                                                 float& x, float& y, float& z) {
  if (a.affine) {
    affine_x(a.aff, i, j, k, x, y, z);
    return;
  }
  coords_from(a, phi[v], phi[v + a.n1], phi[v + 2 * (long long)a.n1], i, j, k, x, y, z);
}

__device__ __forceinline__ void sample_coords(const KArgs& a, const float* phi, int v, float& x, float& y, float& z) {
  int i = 0, j = 0, k = 0;
  if (a.affine || a.disp) {
    const int jk = v / a.d1x;
    i = v - jk * a.d1x;
    j = jk % a.d1y;
    k = jk / a.d1y;
  }
  sample_coords_at(a, phi, v, i, j, k, x, y, z);
}

// Pull output transform (vscale, identity added in displacement mode).
__device__ __forceinline__ float pull_out(const KArgs& a, float val, int c, int i, int j, int k) {
  val *= a.vscale;
  c += a.chan0;
  if (a.add_id && c < 3) val += (float)((c == 0 ? i : c == 1 ? j : k) + 1);
  return val;
}

// Wave-wide reductions of the device library (DPP lane moves; the result reaches every
// active lane).
extern "C" __device__ __attribute__((const)) int __ockl_wfred_min_i32(int);
extern "C" __device__ __attribute__((const)) int __ockl_wfred_max_i32(int);
extern "C" __device__ __attribute__((const)) float __ockl_wfred_max_f32(float);
extern "C" __device__ __attribute__((const)) float __ockl_wfred_add_f32(float);

template <int DP, int NEED>
__device__ __forceinline__ bool setup(const KArgs& a, const float* phi, int v, Axis<DP>& ax, Axis<DP>& ay,
                                      Axis<DP>& az) {
  float x, y, z;
  sample_coords(a, phi, v, x, y, z);
  if (!voxok(x, y, z, a.ext, a.lx, a.ly, a.lz)) return false;
  make_axis<DP, NEED>(x, a.dpx, a.d0x, 1, a.bx, ax);
  make_axis<DP, NEED>(y, a.dpy, a.d0y, a.d0x, a.by, ay);
  make_axis<DP, NEED>(z, a.dpz, a.d0z, a.d0x * a.d0y, a.bz, az);
  return true;
}

#define PP_TAP(DP, ax, i) if (DP > 0 || (i) < (ax).n)

// The x taps of one row. Trilinear taps that are adjacent in memory (every x knot but the
// folded ones at the faces) come in one 8-byte load, halving the gather instructions.
template <int DP>
__device__ __forceinline__ void row_taps(const float* __restrict__ row, const Axis<DP>& ax,
                                         float (&v)[Axis<DP>::NP]) {
  constexpr int NP = Axis<DP>::NP;
  if constexpr (DP == 2) {
    if (ax.off[1] == ax.off[0] + 1) {
      typedef float pair_t __attribute__((ext_vector_type(2), aligned(4)));
      const pair_t p = *reinterpret_cast<const pair_t*>(row + ax.off[0]);
      v[0] = p.x;
      v[1] = p.y;
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) v[i] = row[ax.off[i]];
}

// Value: sum_ijk wx_i wy_j wz_k f[...]. Contractions run one axis at a time (x within a
// row, y within a plane, then z), in the same order as the LDS-tiled gathers, so both
// kernel families give the same bits for a sample (boundary signs, folded here into the
// weights and there into the staged data, factor out exactly).
template <int DP, bool LEAN = false>
__device__ __forceinline__ float gather_val(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az) {
  constexpr int NP = Axis<DP>::NP;
  constexpr int KU = LEAN ? 1 : NP;  // z-planes unrolled
  float acc = 0.f;
#pragma unroll KU
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
      float p = 0.f;
#pragma unroll
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          float v[NP];
          row_taps<DP>(f + az.off[k] + ay.off[j], ax, v);
          float r = 0.f;
#pragma unroll
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) r += ax.w[i] * v[i];
          p += ay.w[j] * r;
        }
      acc += az.w[k] * p;
    }
  return acc;
}

// Gradient: (d/dx, d/dy, d/dz)
template <int DP, bool LEAN = false>
__device__ __forceinline__ void gather_grad(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az, float& gx, float& gy, float& gz) {
  constexpr int NP = Axis<DP>::NP;
  constexpr int KU = LEAN ? 1 : NP;  // z-planes unrolled
  gx = gy = gz = 0.f;
#pragma unroll KU
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
      float pv = 0.f, px = 0.f, py = 0.f;
#pragma unroll
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          float v[NP];
          row_taps<DP>(f + az.off[k] + ay.off[j], ax, v);
          float r = 0.f, rd = 0.f;
#pragma unroll
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) {
              r += ax.w[i] * v[i];
              rd += ax.dw[i] * v[i];
            }
          pv += ay.w[j] * r;
          px += ay.w[j] * rd;
          py += ay.dw[j] * r;
        }
      gx += az.w[k] * px;
      gy += az.w[k] * py;
      gz += az.dw[k] * pv;
    }
}

// Value and gradient from one footprint
template <int DP, bool LEAN = false>
__device__ __forceinline__ void gather_val_grad(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                                const Axis<DP>& az, float& val, float& gx, float& gy, float& gz) {
  constexpr int NP = Axis<DP>::NP;
  constexpr int KU = LEAN ? 1 : NP;  // z-planes unrolled
  val = gx = gy = gz = 0.f;
#pragma unroll KU
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
      float pv = 0.f, px = 0.f, py = 0.f;
#pragma unroll
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          float v[NP];
          row_taps<DP>(f + az.off[k] + ay.off[j], ax, v);
          float r = 0.f, rd = 0.f;
#pragma unroll
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) {
              r += ax.w[i] * v[i];
              rd += ax.dw[i] * v[i];
            }
          pv += ay.w[j] * r;
          px += ay.w[j] * rd;
          py += ay.dw[j] * r;
        }
      val += az.w[k] * pv;
      gx += az.w[k] * px;
      gy += az.w[k] * py;
      gz += az.dw[k] * pv;
    }
}

// Hessian: xx, yy, zz, xy, xz, yz
template <int DP>
__device__ __forceinline__ void gather_hess(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az, float* h) {
  constexpr int NP = Axis<DP>::NP;
#pragma unroll
  for (int q = 0; q < 6; ++q) h[q] = 0.f;
#pragma unroll
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
      float pw = 0.f, pwd = 0.f, pwh = 0.f, pd = 0.f, pdd = 0.f, ph = 0.f;
#pragma unroll
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          const float* row = f + az.off[k] + ay.off[j];
          float r = 0.f, rd = 0.f, rh = 0.f;
#pragma unroll
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) {
              const float v = row[ax.off[i]];
              r += ax.w[i] * v;
              rd += ax.dw[i] * v;
              rh += ax.hw[i] * v;
            }
          pw += ay.w[j] * r;
          pwd += ay.w[j] * rd;
          pwh += ay.w[j] * rh;
          pd += ay.dw[j] * r;
          pdd += ay.dw[j] * rd;
          ph += ay.hw[j] * r;
        }
      h[0] += az.w[k] * pwh;
      h[1] += az.w[k] * ph;
      h[2] += az.hw[k] * pw;
      h[3] += az.w[k] * pdd;
      h[4] += az.dw[k] * pwd;
      h[5] += az.dw[k] * pd;
    }
}

// ---- reference operation order, no FMA contraction --------------------------------
// Used for samples with a coordinate on the large-coordinate path (|x| >= 2^20), where
// the weights can be huge and cancel exactly in the reference's order
// (artifacts/C/pushpull_dev.cu:7-26, 33-58, 91-116); the fast contraction above would
// leave a rounding residue there.
#pragma clang fp contract(off)
template <int DP>
__device__ __forceinline__ float dot_ref(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                         const Axis<DP>& az, const float* a, const float* b, const float* c) {
  constexpr int NP = Axis<DP>::NP;
  float acc = 0.f;
#pragma unroll 1
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          const float wjj = c[k] * b[j];
          const float* row = f + az.off[k] + ay.off[j];
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) acc = acc + a[i] * wjj * row[ax.off[i]];
        }
    }
  return acc;
}

template <int DP>
__device__ __forceinline__ void grad_ref(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                         const Axis<DP>& az, float& g0, float& g1, float& g2) {
  constexpr int NP = Axis<DP>::NP;
  g0 = g1 = g2 = 0.f;
#pragma unroll 1
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          const float wx = az.w[k] * ay.w[j], wy = az.w[k] * ay.dw[j], wz = az.dw[k] * ay.w[j];
          const float* row = f + az.off[k] + ay.off[j];
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) {
              float t = row[ax.off[i]];
              g0 = g0 + ax.dw[i] * wx * t;
              t = t * ax.w[i];
              g1 = g1 + wy * t;
              g2 = g2 + wz * t;
            }
        }
    }
}
#pragma clang fp contract(fast)

template <int DP>
__device__ __forceinline__ void hess_ref(const float* __restrict__ f, const Axis<DP>& ax, const Axis<DP>& ay,
                                         const Axis<DP>& az, float* h) {
  h[0] = dot_ref<DP>(f, ax, ay, az, ax.hw, ay.w, az.w);
  h[1] = dot_ref<DP>(f, ax, ay, az, ax.w, ay.hw, az.w);
  h[2] = dot_ref<DP>(f, ax, ay, az, ax.w, ay.w, az.hw);
  h[3] = dot_ref<DP>(f, ax, ay, az, ax.dw, ay.dw, az.w);
  h[4] = dot_ref<DP>(f, ax, ay, az, ax.dw, ay.w, az.dw);
  h[5] = dot_ref<DP>(f, ax, ay, az, ax.w, ay.dw, az.dw);
}

template <int DP>
__device__ __forceinline__ bool is_slow(const Axis<DP>& ax, const Axis<DP>& ay, const Axis<DP>& az) {
  return !(ax.fast && ay.fast && az.fast);
}

// Value / gradient / value+gradient / Hessian of one sample, picking the contraction
// LEAN: one z-plane of taps in flight at a time (for rarely taken paths that must not
// set a kernel's register budget).
template <int DP, bool LEAN = false>
__device__ __forceinline__ float sample_val(const float* f, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az) {
  if (is_slow(ax, ay, az)) return dot_ref<DP>(f, ax, ay, az, ax.w, ay.w, az.w);
  return gather_val<DP, LEAN>(f, ax, ay, az);
}
template <int DP, bool LEAN = false>
__device__ __forceinline__ void sample_grad(const float* f, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az, float& gx, float& gy, float& gz) {
  if (is_slow(ax, ay, az)) grad_ref<DP>(f, ax, ay, az, gx, gy, gz);
  else gather_grad<DP, LEAN>(f, ax, ay, az, gx, gy, gz);
}
template <int DP, bool LEAN = false>
__device__ __forceinline__ void sample_val_grad(const float* f, const Axis<DP>& ax, const Axis<DP>& ay,
                                                const Axis<DP>& az, float& v, float& gx, float& gy, float& gz) {
  if (is_slow(ax, ay, az)) {
    v = dot_ref<DP>(f, ax, ay, az, ax.w, ay.w, az.w);
    grad_ref<DP>(f, ax, ay, az, gx, gy, gz);
  } else {
    gather_val_grad<DP, LEAN>(f, ax, ay, az, v, gx, gy, gz);
  }
}
template <int DP>
__device__ __forceinline__ void sample_hess(const float* f, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az, float* h) {
  if (is_slow(ax, ay, az)) hess_ref<DP>(f, ax, ay, az, h);
  else gather_hess<DP>(f, ax, ay, az, h);
}

// Scatter: dst[...] += (wx_i * (cy_j*cz_k)) * val  -- wy/wz are passed per op
template <int DP>
__device__ __forceinline__ void scatter_val(float* __restrict__ dst, const Axis<DP>& ax, const Axis<DP>& ay,
                                            const Axis<DP>& az, float val) {
  constexpr int NP = Axis<DP>::NP;
#pragma unroll
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
#pragma unroll
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          float* row = dst + az.off[k] + ay.off[j];
          const float wv = (az.w[k] * ay.w[j]) * val;
#pragma unroll
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i) unsafeAtomicAdd(row + ax.off[i], ax.w[i] * wv);
        }
    }
}

// Scatter of a gradient triple: weight = g0 dwx wy wz + g1 wx dwy wz + g2 wx wy dwz
template <int DP>
__device__ __forceinline__ void scatter_grad(float* __restrict__ dst, const Axis<DP>& ax, const Axis<DP>& ay,
                                             const Axis<DP>& az, float g0, float g1, float g2) {
  constexpr int NP = Axis<DP>::NP;
#pragma unroll
  for (int k = 0; k < NP; ++k) PP_TAP(DP, az, k) {
#pragma unroll
      for (int j = 0; j < NP; ++j) PP_TAP(DP, ay, j) {
          float* row = dst + az.off[k] + ay.off[j];
          const float cx = (az.w[k] * ay.w[j]) * g0;
          const float cyz = (az.w[k] * ay.dw[j]) * g1 + (az.dw[k] * ay.w[j]) * g2;
#pragma unroll
          for (int i = 0; i < NP; ++i) PP_TAP(DP, ax, i)
            unsafeAtomicAdd(row + ax.off[i], ax.dw[i] * cx + ax.w[i] * cyz);
        }
    }
}

// Write (or add, for the later channel groups of a channel-reducing op) a 3-vector.
__device__ __forceinline__ void put3(float* o, long long n1, float a0, float a1, float a2, int accum) {
  if (accum) { a0 += o[0]; a1 += o[n1]; a2 += o[2 * n1]; }
  o[0] = a0; o[n1] = a1; o[2 * n1] = a2;
}

}  // namespace pp
