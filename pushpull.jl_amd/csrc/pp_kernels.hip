// pp_kernels.hip -- gfx950 kernels of the pull/push path.
//
// One work-item per voxel of the phi grid. Per sample the three axes' weights and
// folded offsets live in registers (pp_bspline.hpp) and are reused across all channels
// of the launch. Gathers contract separably: each row of dp taps along x is reduced
// with the x weights first, then weighted by w_y*w_z. Scatters use the hardware float
// atomic (global_atomic_add_f32). Block ids are remapped so that each XCD's L2 sees a
// contiguous range of voxels (neighbouring rows share most of their source footprint).
#define PP_KERNEL_SOURCE 1  // kernel arguments carry the global address space (pp_kernels.hpp)
#include "pp_sample.hpp"

namespace pp {

#define PP_KERNEL_PROLOGUE                                       \
  const int bb = blockIdx.y;                                       \
  const int v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; \
  if (v >= a.n1) return;                                           \
  const float* phi = a.phi + bb * a.phi_bs;

// r = c ? p : q, field by field (selects in registers; a reference chosen at run time
// would put both axes in scratch memory).
template <int DP>
__device__ __forceinline__ void pick_axis(bool c, const Axis<DP>& p, const Axis<DP>& q, Axis<DP>& r) {
#pragma unroll
  for (int i = 0; i < Axis<DP>::NP; ++i) {
    r.w[i] = c ? p.w[i] : q.w[i];
    r.off[i] = c ? p.off[i] : q.off[i];
  }
  r.n = p.n;
  r.fast = p.fast;
}

// Pull of channels whose boundary codes differ only between reflect (1) and antisymmetric
// (2) per axis (csign): the two codes fold to the same indices, so each axis is built with
// and without the sign once and every channel takes its own combination.
template <int DP>
__global__ __launch_bounds__(256) void k_pull_signed(KArgs a) {
  PP_KERNEL_PROLOGUE
  float* out = a.out + bb * a.out_bs + v;
  float x, y, z;
  sample_coords(a, phi, v, x, y, z);
  if (!voxok(x, y, z, a.ext, a.lx, a.ly, a.lz)) {
    for (int c = 0; c < a.nchan; ++c) out[c * a.out_cs] = __builtin_nanf("");
    return;
  }
  // per axis: with the launch's code (s*), and with code 1 where that is 2 (u*)
  Axis<DP> sx, sy, sz, ux, uy, uz;
  make_axis<DP, kW>(x, a.dpx, a.d0x, 1, a.bx, sx);
  make_axis<DP, kW>(y, a.dpy, a.d0y, a.d0x, a.by, sy);
  make_axis<DP, kW>(z, a.dpz, a.d0z, a.d0x * a.d0y, a.bz, sz);
  make_axis<DP, kW>(x, a.dpx, a.d0x, 1, a.bx == 2 ? 1 : a.bx, ux);
  make_axis<DP, kW>(y, a.dpy, a.d0y, a.d0x, a.by == 2 ? 1 : a.by, uy);
  make_axis<DP, kW>(z, a.dpz, a.d0z, a.d0x * a.d0y, a.bz == 2 ? 1 : a.bz, uz);
  const float* f = a.f + bb * a.f_bs;
  const int jk = v / a.d1x, i = v - jk * a.d1x;
  for (int c = 0; c < a.nchan; ++c) {
    const unsigned sg = chan_signs(a, c);
    Axis<DP> ax, ay, az;
    pick_axis((sg & 1u) != 0, sx, ux, ax);
    pick_axis((sg & 2u) != 0, sy, uy, ay);
    pick_axis((sg & 4u) != 0, sz, uz, az);
    out[c * a.out_cs] = pull_out(a, sample_val<DP>(f + c * a.f_cs, ax, ay, az), c, i, jk % a.d1y, jk / a.d1y);
  }
}

template <int DP>
__global__ __launch_bounds__(256) void k_pull(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  float* out = a.out + bb * a.out_bs + v;
  if (!setup<DP, kW>(a, phi, v, ax, ay, az)) {
    for (int c = 0; c < a.nchan; ++c) out[c * a.out_cs] = __builtin_nanf("");
    return;
  }
  const float* f = a.f + bb * a.f_bs;
  if (a.vscale != 1.0f || a.add_id) {
    const int jk = v / a.d1x, i = v - jk * a.d1x;
    for (int c = 0; c < a.nchan; ++c)
      out[c * a.out_cs] = pull_out(a, sample_val<DP>(f + c * a.f_cs, ax, ay, az), c, i, jk % a.d1y, jk / a.d1y);
    return;
  }
  for (int c = 0; c < a.nchan; ++c) out[c * a.out_cs] = sample_val<DP>(f + c * a.f_cs, ax, ay, az);
}

template <int DP>
__global__ __launch_bounds__(256) void k_pullg(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  float* out = a.out + bb * a.out_bs + v;
  const long long n1 = a.n1;
  if (!setup<DP, kW | kDW>(a, phi, v, ax, ay, az)) {
    const float nan = __builtin_nanf("");
    for (int c = 0; c < a.nchan; ++c) {
      float* o = out + c * a.out_cs;
      o[0] = nan; o[n1] = nan; o[2 * n1] = nan;
    }
    return;
  }
  const float* f = a.f + bb * a.f_bs;
  for (int c = 0; c < a.nchan; ++c) {
    float gx, gy, gz;
    sample_grad<DP>(f + c * a.f_cs, ax, ay, az, gx, gy, gz);
    float* o = out + c * a.out_cs;
    o[0] = gx; o[n1] = gy; o[2 * n1] = gz;
  }
}

template <int DP>
__global__ __launch_bounds__(256) void k_pullh(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  float* out = a.out + bb * a.out_bs + v;
  const long long n1 = a.n1;
  if (!setup<DP, kW | kDW | kHW>(a, phi, v, ax, ay, az)) {
    const float nan = __builtin_nanf("");
    for (int c = 0; c < a.nchan; ++c)
      for (int q = 0; q < 9; ++q) out[c * a.out_cs + q * n1] = nan;
    return;
  }
  const float* f = a.f + bb * a.f_bs;
  for (int c = 0; c < a.nchan; ++c) {
    float h[6];
    sample_hess<DP>(f + c * a.f_cs, ax, ay, az, h);
    float* o = out + c * a.out_cs;
    // column-major 3x3 (pushpull.c:88-93)
    o[0] = h[0]; o[4 * n1] = h[1]; o[8 * n1] = h[2];
    o[1 * n1] = h[3]; o[3 * n1] = h[3];
    o[2 * n1] = h[4]; o[6 * n1] = h[4];
    o[5 * n1] = h[5]; o[7 * n1] = h[5];
  }
}

template <int DP>
__global__ __launch_bounds__(256) void k_push(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  const float* f1 = a.f + bb * a.f_bs + v;
  if (!setup<DP, kW>(a, phi, v, ax, ay, az)) return;
  float* dst = a.out + bb * a.out_bs;
  int idx[3] = {0, 0, 0};
  if (a.vmode == 2) {
    const int jk = v / a.d1x;
    idx[0] = v - jk * a.d1x;
    idx[1] = jk % a.d1y;
    idx[2] = jk / a.d1y;
  }
  for (int c = 0; c < a.nchan; ++c) {
    const int ch = c + a.chan0;
    const float val = a.vmode == 0 ? f1[c * a.f_cs] : a.vmode == 1 ? 1.0f : (float)(idx[ch < 3 ? ch : 0] + 1);
    if (isfinite(val)) scatter_val<DP>(dst + c * a.out_cs, ax, ay, az, val);
  }
}

template <int DP>
__global__ __launch_bounds__(256) void k_pushg(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  if (!setup<DP, kW | kDW>(a, phi, v, ax, ay, az)) return;
  const float* g1 = a.f + bb * a.f_bs + v;
  const long long n1 = a.n1;
  float* dst = a.out + bb * a.out_bs;
  for (int c = 0; c < a.nchan; ++c) {
    const float* g = g1 + c * a.f_cs;
    scatter_grad<DP>(dst + c * a.out_cs, ax, ay, az, g[0], g[n1], g[2 * n1]);
  }
}

// dphi_a = sum_c delta_c * d f_c / d x_a   (NaN where the sample is invalid)
template <int DP>
__global__ __launch_bounds__(256) void k_pull_bwd(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  float* out = a.out + bb * a.out_bs + v;
  const long long n1 = a.n1;
  if (!setup<DP, kW | kDW>(a, phi, v, ax, ay, az)) {
    const float nan = __builtin_nanf("");
    out[0] = nan; out[n1] = nan; out[2 * n1] = nan;
    return;
  }
  const float* f = a.f + bb * a.f_bs;
  const float* d = a.aux + bb * a.aux_bs + v;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int c = 0; c < a.nchan; ++c) {
    float gx, gy, gz;
    sample_grad<DP>(f + c * a.f_cs, ax, ay, az, gx, gy, gz);
    const float dc = d[c * a.aux_cs];
    sx += dc * gx; sy += dc * gy; sz += dc * gz;
  }
  put3(out, n1, sx, sy, sz, a.accum);
}

// dphi_b = sum_c sum_a delta_{a,c} H_{a,b,c}
template <int DP>
__global__ __launch_bounds__(256) void k_pullg_bwd(KArgs a) {
  PP_KERNEL_PROLOGUE
  Axis<DP> ax, ay, az;
  float* out = a.out + bb * a.out_bs + v;
  const long long n1 = a.n1;
  if (!setup<DP, kW | kDW | kHW>(a, phi, v, ax, ay, az)) {
    const float nan = __builtin_nanf("");
    out[0] = nan; out[n1] = nan; out[2 * n1] = nan;
    return;
  }
  const float* f = a.f + bb * a.f_bs;
  const float* d = a.aux + bb * a.aux_bs + v;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  for (int c = 0; c < a.nchan; ++c) {
    float h[6];
    sample_hess<DP>(f + c * a.f_cs, ax, ay, az, h);
    const float* dc = d + c * a.aux_cs;
    const float d0 = dc[0], d1 = dc[n1], d2 = dc[2 * n1];
    // H = [xx xy xz; xy yy yz; xz yz zz]
    s0 += d0 * h[0] + d1 * h[3] + d2 * h[4];
    s1 += d0 * h[3] + d1 * h[1] + d2 * h[5];
    s2 += d0 * h[4] + d1 * h[5] + d2 * h[2];
  }
  put3(out, n1, s0, s1, s2, a.accum);
}

// Fused matching objective: r_c = pull(f)_c - g_c; loss += r_c^2/2 over finite r;
// dphi = sum_c r_c grad f_c (NaN where invalid); df0 += push(r) (skips non-finite r).
template <int DP>
__global__ __launch_bounds__(256) void k_objgrad(KArgs a) {
  const int bb = blockIdx.y;
  const int v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  float lsum = 0.f;
  if (v < a.n1) {
    const float* phi = a.phi + bb * a.phi_bs;
    Axis<DP> ax, ay, az;
    float* out = a.out + bb * a.out_bs + v;
    const long long n1 = a.n1;
    if (!setup<DP, kW | kDW>(a, phi, v, ax, ay, az)) {
      const float nan = __builtin_nanf("");
      out[0] = nan; out[n1] = nan; out[2 * n1] = nan;
    } else {
      const float* f = a.f + bb * a.f_bs;
      const float* g = a.aux + bb * a.aux_bs + v;
      float* df = a.out2 + bb * a.out2_bs;
      float sx = 0.f, sy = 0.f, sz = 0.f;
      for (int c = 0; c < a.nchan; ++c) {
        float val, gx, gy, gz;
        sample_val_grad<DP>(f + c * a.f_cs, ax, ay, az, val, gx, gy, gz);
        const float r = val - g[c * a.aux_cs];
        sx += r * gx; sy += r * gy; sz += r * gz;
        if (isfinite(r)) {
          lsum += 0.5f * r * r;
          scatter_val<DP>(df + c * a.out2_cs, ax, ay, az, r);
        }
      }
      put3(out, n1, sx, sy, sz, a.accum);
    }
  }
  // block reduction of the loss, one double atomic per block
  lsum = __ockl_wfred_add_f32(lsum);
  __shared__ float wsum[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += (double)wsum[w];
    unsafeAtomicAdd(a.red, s);
  }
}

// Vector-Jacobian product of pull (both cotangents of the rrule, src/pushpull_pb.jl:11-17):
// dphi = sum_c delta_c grad f_c (NaN where invalid); df0 += push(delta) (skips non-finite
// delta, pushpull.c:114).
template <int DP>
__global__ __launch_bounds__(256) void k_pull_vjp(KArgs a) {
  const int bb = blockIdx.y;
  const int v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (v >= a.n1) return;
  const float* phi = a.phi + bb * a.phi_bs;
  Axis<DP> ax, ay, az;
  float* out = a.out + bb * a.out_bs + v;
  const long long n1 = a.n1;
  if (!setup<DP, kW | kDW>(a, phi, v, ax, ay, az)) {
    const float nan = __builtin_nanf("");
    out[0] = nan; out[n1] = nan; out[2 * n1] = nan;
    return;
  }
  const float* f = a.f + bb * a.f_bs;
  const float* d = a.aux + bb * a.aux_bs + v;
  float* df = a.out2 + bb * a.out2_bs;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int c = 0; c < a.nchan; ++c) {
    float gx, gy, gz;
    sample_grad<DP>(f + c * a.f_cs, ax, ay, az, gx, gy, gz);
    const float w = d[c * a.aux_cs];
    sx += w * gx; sy += w * gy; sz += w * gz;
    if (isfinite(w)) scatter_val<DP>(df + c * a.out2_cs, ax, ay, az, w);
  }
  put3(out, n1, sx, sy, sz, a.accum);
}

// p[0..n) = v with 16-byte stores: scalar head up to the first 16-byte boundary, float4
// body, scalar tail.
__global__ void k_fill(float* p, long long n, float v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long head = (long long)(((16u - ((unsigned)(uintptr_t)p & 15u)) & 15u) / 4u);
  if (head > n) head = n;
  if (t < head) p[t] = v;
  float4* q = reinterpret_cast<float4*>(p + head);
  const long long nq = (n - head) / 4;
  const float4 vv = make_float4(v, v, v, v);
  for (long long i = t; i < nq; i += stride) q[i] = vv;
  const long long done = head + 4 * nq;
  if (t < n - done) p[done + t] = v;
}

__global__ void k_axpy(float* __restrict__ y, const float* __restrict__ x, float a, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = fmaf(a, x[i], y[i]);  // (the fused objective step rounds alike)
}

__global__ void k_fill_d(double* p, long long n, double v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------------
template <int DP>
static hipError_t launch_dp(Op op, const KArgs& a, int nbatch, hipStream_t s, int block) {
  const dim3 grid((unsigned)((a.n1 + block - 1) / block), (unsigned)nbatch, 1);
  switch (op) {
    case OP_PULL:
      if (a.csign_on) hipLaunchKernelGGL(k_pull_signed<DP>, grid, dim3(block), 0, s, a);
      else hipLaunchKernelGGL(k_pull<DP>, grid, dim3(block), 0, s, a);
      break;
    case OP_PULLG: hipLaunchKernelGGL(k_pullg<DP>, grid, dim3(block), 0, s, a); break;
    case OP_PULLH: hipLaunchKernelGGL(k_pullh<DP>, grid, dim3(block), 0, s, a); break;
    case OP_PUSH: hipLaunchKernelGGL(k_push<DP>, grid, dim3(block), 0, s, a); break;
    case OP_PUSHG: hipLaunchKernelGGL(k_pushg<DP>, grid, dim3(block), 0, s, a); break;
    case OP_PULL_BWD: hipLaunchKernelGGL(k_pull_bwd<DP>, grid, dim3(block), 0, s, a); break;
    case OP_PULLG_BWD: hipLaunchKernelGGL(k_pullg_bwd<DP>, grid, dim3(block), 0, s, a); break;
    case OP_OBJGRAD: hipLaunchKernelGGL(k_objgrad<DP>, grid, dim3(block), 0, s, a); break;
    case OP_PULL_VJP: hipLaunchKernelGGL(k_pull_vjp<DP>, grid, dim3(block), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_op(Op op, const KArgs& a, int nbatch, hipStream_t s, int block) {
  if (a.n1 <= 0 || nbatch <= 0) return hipSuccess;
  const bool iso = a.dpx == a.dpy && a.dpx == a.dpz;
  if (!iso) return launch_dp<0>(op, a, nbatch, s, block);
  switch (a.dpx) {
    case 1: return launch_dp<1>(op, a, nbatch, s, block);
    case 2: return launch_dp<2>(op, a, nbatch, s, block);
    case 3: return launch_dp<3>(op, a, nbatch, s, block);
    case 4: return launch_dp<4>(op, a, nbatch, s, block);
    case 5: return launch_dp<5>(op, a, nbatch, s, block);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fill(float* p, long long n, float v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long nb = (n / 4 + 255) / 256 + 1;  // a float4 per thread, and the head / tail
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_axpy(float* y, const float* x, float a, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)nb), dim3(256), 0, s, y, x, a, n);
  return hipGetLastError();
}

hipError_t launch_fill_d(double* p, long long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill_d, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace pp
