// pp_tiled.hip -- tiled gfx950 kernels for the hot ops (pull, pull_grad, pull_hess, push,
// push_grad, the phi pullbacks, the pull VJP and the fused matching objective): launchers
// for isotropic degrees, the batched launch over uneven volumes and the LDS-atomic
// calibration kernel. The device code is in pp_tiled_impl.hpp:
//
// A 256-thread workgroup owns a TX x TY x TZ tile of the phi grid (TX*TY*TZ/256 samples
// per thread). Its in-window samples' coordinates are reduced to a bounding box of knots
// in the *unfolded* index space of the image grid; the box lives in LDS:
//   gathers  stage the box once (row loads, boundary fold and sign applied while
//            staging), then every sample reads its dp^3 taps from LDS;
//   scatters accumulate every tap in 64-bit fixed point with an LDS integer atomic, then
//            flush each non-zero cell with one folded global float atomic -- a few global
//            atomics per sample instead of dp^3.
// Samples whose coordinates lie outside a window around the tile's centre sample
// (displaced far away, sent outside the field of view, or on the large-coordinate path)
// are handled individually through global memory, so a few outliers never inflate the
// box; a box beyond the launch's LDS budget is clipped to it, and the samples that leave
// it take the same global path.
#include "pp_tiled_impl.hpp"

namespace pp {

// ---- launcher ------------------------------------------------------------------
// Box budgets (floats of dynamic LDS per workgroup), per op and degree: small boxes let
// more workgroups share a CU, whose latency chains (load phi, size the box, stage,
// compute) then overlap; a box beyond the budget is clipped to it and the samples that
// leave it run per sample through global memory. Each default is the largest budget that
// keeps the workgroups per CU its registers allow (160 KiB of LDS per CU, 224 bytes of
// static LDS per workgroup): gathers stage floats; scatters, the VJP and the objective
// accumulate 64-bit cells (two floats each); the row table takes two floats a box row.
// A caller may pass its own budget (the autotuner does).
int default_cap(Op op, int dp) {
  // (workgroups per CU by LDS: 6 up to 6770 floats, 5 up to 8136, 4 up to 10184, 3 up to 13596)
  const int d = dp < 2 ? 2 : dp > 5 ? 5 : dp;
  switch (op) {
    case OP_PULL:
    case OP_PULLG:
    case OP_PULLH: return d == 2 ? 4608 : d == 3 ? 6770 : d == 4 ? 8136 : 10184;
    case OP_PULL_BWD:
    case OP_PULLG_BWD: return d == 2 ? 4608 : d == 3 ? 8136 : 10184;
    case OP_PUSH:
    case OP_PUSHG: return d <= 3 ? 8136 : d == 4 ? 10184 : 13596;  // 64-bit cells
    default: return d <= 4 ? 10184 : 13596;                         // the VJP and the objective
  }
}

template <int DP>
static hipError_t launch_deg(Op op, const KArgs& a, int nbatch, hipStream_t s, TileCfg cfg) {
  const int cap = cfg.cap > 0 ? cfg.cap : default_cap(op, DP);
  switch (cfg.shape) {
#ifndef PP_DEFAULT_SHAPE_ONLY
    case 1: return launch_shape<DP, 32, 8, 4>(op, a, nbatch, s, cap);
    case 2: return launch_shape<DP, 16, 16, 4>(op, a, nbatch, s, cap);
#endif
    default: return launch_shape<DP, 16, 8, 8>(op, a, nbatch, s, cap);
  }
}

int batch_tiles(const KArgs& a) { return tiles_of<kBTX, kBTY, kBTZ>(a); }

// ---- LDS atomic calibration --------------------------------------------------------
// The rate of LDS integer adds on cells of type T (the scatters' 64-bit cells) under the
// scatter's own access pattern, the on-chip bound of push and the VJP: every lane adds the
// 4x4x4 taps of cubic samples into a box shaped like
// a 16x8x8 tile's at degree 3 (20 x 12 x 12 cells), its knots following the lane's x and y
// (z: 2 per repetition, as a thread's samples), a third of the lanes shifted by a cell
// along x and a sixth along y as under a smooth displacement. `iters` samples per lane, one
// atomic add per tap; out (gridDim.x * 256 words) receives checksums of the box.
template <class T>
__global__ __launch_bounds__(kTB) void k_lds_add_rate(acc_t* __restrict__ out, int iters) {
  constexpr int BXc = 20, BYc = 12, BZc = 12, CELLS = BXc * BYc * BZc;
  __shared__ T cal[CELLS];
  for (int i = threadIdx.x; i < CELLS; i += kTB) cal[i] = 0;
  __syncthreads();
  const int l = threadIdx.x;
  const unsigned h = (blockIdx.x * kTB + l) * 2654435761u;
  const int jx = ((h >> 24) & 255u) < 85u ? 1 : 0, jy = ((h >> 16) & 255u) < 43u ? 1 : 0;
  // largest cell: 16 + 20 * 8 + 240 * (2 + 6) + 3 + 20 * 3 + 240 * 3 = 2879 < CELLS
  const int base = (l % 16 + jx) + BXc * ((l / 16) % 8 + jy) + BXc * BYc * (2 * (l / 128));
  for (int it = 0; it < iters; ++it) {
    const int b = base + BXc * BYc * 2 * (it & 3);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) atomicAdd(cal + b + i + BXc * j + BXc * BYc * k, (T)(it + i + 1));
  }
  __syncthreads();
  T sum = 0;
  for (int i = l; i < CELLS; i += kTB) sum += cal[i];
  out[(size_t)blockIdx.x * kTB + l] = (acc_t)sum;
}

hipError_t launch_lds_add_rate(unsigned long long* out, int blocks, int iters, hipStream_t s) {
  hipLaunchKernelGGL((k_lds_add_rate<acc_t>), dim3((unsigned)blocks), dim3(kTB),
                     0, s, out, iters);
  return hipGetLastError();
}

// Zero what a batched push / objective accumulates into (graph-capturable).
__global__ void k_batch_zero(const BatchVol* __restrict__ vols, int kind) {
  const BatchVol& b = vols[blockIdx.y];
  PP_GLOBAL float* p = (kind == OP_OBJGRAD || kind == OP_PULL_VJP) ? b.a.out2 : b.a.out;
  const long long n = b.zero_n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
  if (kind == OP_OBJGRAD && blockIdx.x == 0 && threadIdx.x == 0) b.a.red[0] = 0.0;
}


hipError_t launch_batch_zero(Op op, const BatchVol* dvols, int nvol, hipStream_t s) {
  if (nvol <= 0 || !(op == OP_PUSH || op == OP_PUSHG || op == OP_OBJGRAD || op == OP_PULL_VJP)) return hipSuccess;
  hipLaunchKernelGGL(k_batch_zero, dim3(512, (unsigned)nvol), dim3(256), 0, s, dvols, (int)op);
  return hipGetLastError();
}

hipError_t launch_batch_group(Op op, int dp, bool an, const BatchVol* dvols, int nvol, const XcdSplit& split,
                              hipStream_t s, int cap_override) {
  int longest = 0;
  for (int x = 0; x < kBatchXcds; ++x) longest = std::max(longest, split.s[x + 1] - split.s[x]);
  if (nvol <= 0 || longest <= 0) return hipSuccess;
  const void* k = an ? batch_kernel_aniso(op, dp) : batch_kernel_for<false>(op, dp);
  if (!k) return hipErrorNotSupported;
  XcdSplit sp = split;
  int cap = cap_override > 0 ? cap_override : default_cap(op, dp);
  if (dp < 2) cap = 2;  // degree 0 runs per sample (no box)
  void* args[] = {(void*)&dvols, (void*)&nvol, (void*)&sp, (void*)&cap};
  const long long grid = (long long)longest * kBatchXcds;
  if (grid > INT_MAX) return hipErrorInvalidValue;
  return hipLaunchKernel(k, dim3((unsigned)grid), dim3(kTB), args, cap_bytes(cap), s);
}

hipError_t launch_tiled(Op op, const KArgs& a, int nbatch, hipStream_t s, TileCfg cfg) {
  if (op != OP_PULL && op != OP_PULLG && op != OP_PULLH && op != OP_PULL_BWD && op != OP_PULLG_BWD &&
      op != OP_OBJGRAD && op != OP_PULL_VJP && op != OP_PUSH && op != OP_PUSHG)
    return hipErrorNotSupported;
  if (a.n1 <= 0 || nbatch <= 0) return hipSuccess;
  if ((long long)a.d1x * a.d1y * a.d1z != a.n1) return hipErrorNotSupported;
  if (a.dpx != a.dpy || a.dpx != a.dpz) return launch_tiled_aniso(op, a, nbatch, s, cfg);
  switch (a.dpx) {
    case 2: return launch_deg<2>(op, a, nbatch, s, cfg);
    case 3: return launch_deg<3>(op, a, nbatch, s, cfg);
    case 4: return launch_deg<4>(op, a, nbatch, s, cfg);
    case 5: return launch_deg<5>(op, a, nbatch, s, cfg);
    default: return hipErrorNotSupported;
  }
}

}  // namespace pp
