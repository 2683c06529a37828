// pp_kernels.hpp -- host-side view of the gfx950 kernels (internal to libpushpull_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

// Operations on the path. Names follow artifacts/C/pushpull.c / src/pushpull_pb.jl.
enum Op : int {
  OP_PULL = 0,      // f1 = pull(f0, phi)                          pushpull.c:14
  OP_PULLG = 1,     // g1 = pull_grad(f0, phi)        (3 comps)    pushpull.c:37
  OP_PULLH = 2,     // h1 = pull_hess(f0, phi)        (9 comps)    pushpull.c:70
  OP_PUSH = 3,      // f0 += push(f1, phi)                         pushpull.c:106
  OP_PUSHG = 4,     // f0 += push_grad(g1, phi)                    pushpull.cu:145
  OP_PULL_BWD = 5,  // dphi = sum_c delta_c grad f_c               pushpull_pb.jl:13-17
  OP_PULLG_BWD = 6, // dphi = sum_{a,c} delta_{a,c} H_{a,.,c}      pushpull_pb.jl:26-35
  OP_OBJGRAD = 7,   // r = pull(f0)-g; dphi = sum_c r_c grad f_c; df0 += push(r); loss += r^2/2
  OP_PULL_VJP = 8,  // dphi = sum_c delta_c grad f_c; df0 += push(delta)  pushpull_pb.jl:11-17
  OP_COUNT = 9
};

// Every array in the kernel arguments lives in device global memory. In the kernel
// sources' device pass the pointers carry that address space, so that pointers read from
// a table of arguments (the batched launches) still compile to global rather than flat
// memory instructions; the layout is the same in every pass.
#if defined(__HIP_DEVICE_COMPILE__) && defined(PP_KERNEL_SOURCE)
#define PP_GLOBAL __attribute__((address_space(1)))
#else
#define PP_GLOBAL
#endif

constexpr int kXcds = 8;

// Round-robin dispatch puts block b on XCD b % 8; give XCD k a contiguous chunk of the
// nb blocks, so that neighbouring blocks share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb / kXcds, rem = nb % kXcds;
  const int x = b % kXcds, i = b / kXcds;
  return (x < rem) ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// Kernel arguments, passed by value (kernarg segment; no module globals, so launches
// on different streams never share state).
struct KArgs {
  const PP_GLOBAL float* f;    // gathered image: f0 (pull family) / pushed values (push family)
  const PP_GLOBAL float* phi;  // deformation, 1-based coordinates, 3 comps of n1 voxels
  const PP_GLOBAL float* aux;  // upstream gradient (backward ops) / target image (objgrad)
  PP_GLOBAL float* out;        // output image / dphi
  PP_GLOBAL float* out2;       // objgrad: dL/df0 (scatter target)
  PP_GLOBAL double* red;       // objgrad: loss accumulator (one double)
  long long f_cs, f_bs;      // channel / batch strides of f (elements)
  long long phi_bs;          // batch stride of phi
  long long aux_cs, aux_bs;  // strides of aux
  long long out_cs, out_bs;  // strides of out
  long long out2_cs, out2_bs;
  int n1;                    // voxels in the phi grid
  int d1x, d1y, d1z;         // phi grid (tiled kernels; n1 == d1x*d1y*d1z)
  int d0x, d0y, d0z;         // grid sampled by pull / scattered into by push
  int dpx, dpy, dpz;         // dp = degree + 1 per axis (used by the runtime-degree kernels)
  int bx, by, bz;            // boundary code per axis
  int ext;                   // extrapolate flag
  int nchan;                 // channels processed per thread
  int accum;                 // channel-reducing ops: add to `out` instead of overwriting
  float lx, ly, lz;          // d0 - 0.99f per axis (VOXOK upper bounds)
  int affine;                // 1: sample coordinates come from aff, not from phi
  float aff[12];             // 0-based 3x4 affine, column-major (artifacts/C/pushpull.c:185-201)
  // Per-channel boundary signs (the tiled kernels): all channels fold with bx/by/bz;
  // channel c takes the antisymmetric sign along axis a only where bit 3c+a of csign is
  // set (codes 1 and 2 fold to the same index and differ only in sign).
  int csign_on;
  unsigned csign;
  // Displacement mode (scaling and squaring): phi holds a displacement u and the sample's
  // 1-based coordinate is (index + 1) + dscale * u, rounded as the sum of two floats.
  int disp;
  float dscale;
  // Pull outputs: value * vscale, plus the 1-based index along axis c for channel c < 3
  // when add_id is set.
  float vscale;
  int add_id;
  int chan0;  // channel index of the launch's first channel (per-channel launches)
  // Push values: 0 from f; 1 ones (a count image); 2 the 1-based index of the sample along
  // axis (channel + chan0) (pushing the identity, src/multigrid.jl:270). With out2 set, the
  // tiled push also scatters ones into out2 as one more channel (index nchan): the count.
  int vmode;
  // Objective (tiled): when set, phi_next[v] = fmaf(step, dphi[v], phi[v]) for every phi
  // voxel v of the tile once its dphi is final (a descent step fused into the gradient
  // pass; may be phi itself, since each voxel's phi is read only by the tile that owns it).
  PP_GLOBAL float* phi_next;
  float step;
};

// Boundary code of channel c along `axis` for a launch folding with `code`.
__host__ __device__ inline int chan_code(const KArgs& a, int c, int axis, int code) {
  return (a.csign_on && code == 2 && !((a.csign >> (3 * c + axis)) & 1u)) ? 1 : code;
}
// Axes (bit a) along which channel c takes the boundary sign.
__host__ __device__ inline unsigned chan_signs(const KArgs& a, int c) {
  return a.csign_on ? (a.csign >> (3 * c)) & 7u : 7u;
}

// Launch op on `stream`; dp is (DP,DP,DP) when isotropic, else the runtime path is used.
// grid.y = nbatch. Returns hipSuccess or the launch error.
hipError_t launch_op(Op op, const KArgs& a, int nbatch, hipStream_t stream, int block);

// Tiled variants (pp_tiled.hip): one workgroup per tile of the phi grid, source
// footprint staged in / scattered into LDS. Returns hipErrorNotSupported for an op or
// degree they do not cover (the caller then uses launch_op).
struct TileCfg {
  int shape;  // index into the compiled tile shapes (0: 16x8x8, 1: 32x8x4, 2: 16x16x4)
  int cap;    // LDS box budget in floats (0: the default for the op and degree)
};
// Default LDS box budget (floats) of the tiled kernel of op at isotropic dp.
int default_cap(Op op, int dp);
hipError_t launch_tiled(Op op, const KArgs& a, int nbatch, hipStream_t stream, TileCfg cfg);
// ... for degrees that differ between axes (pp_tiled_aniso.hip): the 16x8x8 kernels of the
// largest degree, every axis's taps in that window (cfg.shape is not used).
hipError_t launch_tiled_aniso(Op op, const KArgs& a, int nbatch, hipStream_t stream, TileCfg cfg);

// One volume of a batched launch (device-resident array, built by the batch plan).
struct BatchVol {
  KArgs a;          // one volume, all its channels; pointers already at the volume
  int dp;           // dp = degree + 1 (the largest, for a degree per axis)
  int an;           // 1: the degrees differ between axes (kernels of their own)
  int tile0;        // first tile number of this volume
  long long zero_n; // floats zeroed before a push / objective launch
};
int batch_tiles(const KArgs& a);   // tiles of one volume (16x8x8 tiles of the phi grid)
// Contiguous tile ranges of a degree group, one per XCD: XCD x takes tiles
// [s[x], s[x+1]), chosen so that the ranges carry equal work (a tile costs its volume's
// channel count plus one).
constexpr int kBatchXcds = 8;
struct XcdSplit {
  int s[kBatchXcds + 1];
};
hipError_t launch_batch_zero(Op op, const BatchVol* dvols, int nvol, hipStream_t stream);
// One degree group: dvols[0..nvol) all have this dp; tile0 counts from the group start.
// ... the kernel of a group whose degrees differ between axes (pp_tiled_aniso.hip)
const void* batch_kernel_aniso(Op op, int dp);
hipError_t launch_batch_group(Op op, int dp, bool an, const BatchVol* dvols, int nvol, const XcdSplit& split,
                              hipStream_t stream, int cap_override = 0);

// Fill n floats with v (graph-capturable; no hipMemsetAsync nodes).
hipError_t launch_fill(float* p, long long n, float v, hipStream_t stream);
hipError_t launch_fill_d(double* p, long long n, double v, hipStream_t stream);
hipError_t launch_axpy(float* y, const float* x, float a, long long n, hipStream_t stream);
hipError_t launch_lds_add_rate(unsigned long long* out, int blocks, int iters, hipStream_t stream);

}  // namespace pp
