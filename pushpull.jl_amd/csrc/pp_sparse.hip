// pp_sparse.hip -- gfx950 kernels of the sparse regulariser.
//
// vel2mom (artifacts/C/sparse_operator_dev.cu:164-204): u_i = sum_j (L_ij * v_j), one
// thread per voxel. relax (sparse_operator_dev.cu:89-161): for every voxel of one colour
// lattice, b = g - L v - H v, A = centre of L (+ H), v += chol(A) \ b. Voxels of one
// colour lie dp apart along every axis, so none reads another's value and a launch per
// colour reproduces the sequential sweep of artifacts/C/sparse_operator.c:116-135
// exactly; the host issues the lattices in that sweep's order.
//
// The arithmetic follows the reference operation by operation in FP32: no contraction
// into fused multiply-adds (pragma below), IEEE division and square root, the same
// summation orders. Results are bit-identical to the CPU code.
#define PP_KERNEL_SOURCE 1
#include <hip/hip_runtime.h>

#include "pp_sparse.hpp"

#pragma clang fp contract(off)

namespace pp {
namespace {

constexpr int kSpBlock = 256;

// Fold coordinate c into [0, d) under boundary code `code` (0 circulant, 1 reflect,
// 2 reflect with sign change on the mirrored copies); flips `neg` for a sign change.
// artifacts/C/patch.cu:5-35.
__device__ inline int fold_axis(int c, int d, int code, bool& neg) {
  if ((unsigned)c < (unsigned)d) return c;
  if (code == 0) {
    int m = c % d;
    return m < 0 ? m + d : m;
  }
  const int p = 2 * d;
  int m = c % p;
  if (m < 0) m += p;
  if (m < d) return m;
  neg ^= (code == 2);
  return p - 1 - m;
}

typedef const PP_GLOBAL float* gfloat;

// Value of the stencil entry (code, off) of input channel j (array vj, boundary codes bj)
// around voxel (x, y, z), linear index m.
template <bool EDGE>
__device__ inline float tap(const SpArgs& a, gfloat vj, int code, int off, int x, int y, int z,
                            long long m, const int* bj) {
  if (!EDGE) return vj[m + off];
  const int dx = (code & 255) - 16, dy = ((code >> 8) & 255) - 16, dz = ((code >> 16) & 255) - 16;
  bool neg = false;
  const int ix = fold_axis(x + dx, a.d[0], bj[0], neg);
  const int iy = fold_axis(y + dy, a.d[1], bj[1], neg);
  const int iz = fold_axis(z + dz, a.d[2], bj[2], neg);
  const float f = vj[ix + (long long)a.d[0] * (iy + (long long)a.d[1] * iz)];
  return neg ? -f : f;
}

// sp_conv (sparse_operator_dev.cu:12-20): s = sum_{e>0} w_e (f_e - f0), then
// s += (w0 + sum of the others) * f0.
template <bool EDGE>
__device__ inline float list_conv(const SpArgs& a, int p, gfloat vj, int x, int y, int z,
                                  long long m, const int* bj) {
  const SpPair q = a.pairs[p];
  const SpEntry e0 = a.ent[q.start];
  const float f0 = tap<EDGE>(a, vj, e0.code, e0.off, x, y, z, m, bj);
  float s = 0.0f;
  for (int e = q.start + 1; e < q.start + q.len; ++e) {
    const SpEntry t = a.ent[e];
    s += t.w * (tap<EDGE>(a, vj, t.code, t.off, x, y, z, m, bj) - f0);
  }
  s += q.c0w * f0;
  return s;
}

template <int NC>
__device__ inline void chol_solve(float (&A)[NC * NC], float (&b)[NC], float (&x)[NC]) {
  float p[NC];
  // choldcf (artifacts/C/chol.cu:11-30)
#pragma unroll
  for (int i = 0; i < NC; ++i) {
#pragma unroll
    for (int j = i; j < NC; ++j) {
      float t = A[i * NC + j];
#pragma unroll
      for (int k = i - 1; k >= 0; --k) t -= A[i * NC + k] * A[j * NC + k];
      if (j == i)
        p[i] = sqrtf(t);
      else
        A[j * NC + i] = t / p[i];
    }
  }
  // chollsf (chol.cu:41-59)
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    float t = b[i];
#pragma unroll
    for (int k = i - 1; k >= 0; --k) t -= A[i * NC + k] * x[k];
    x[i] = t / p[i];
  }
#pragma unroll
  for (int i = NC - 1; i >= 0; --i) {
    float t = x[i];
#pragma unroll
    for (int k = i + 1; k < NC; ++k) t -= A[k * NC + i] * x[k];
    x[i] = t / p[i];
  }
}

template <int NC, bool EDGE, int OP>
__global__ __launch_bounds__(kSpBlock) void k_sparse(SpArgs a, SpLattice l) {
  const long long t = (long long)blockIdx.x * kSpBlock + threadIdx.x;
  if (t >= l.n) return;
  // lattice coordinates: q = t / cnt0 and iz = q / cnt1, exact after one correction
  long long q = (long long)((double)t * l.rx);
  if (q * l.cnt[0] > t) --q;
  else if ((q + 1) * l.cnt[0] <= t) ++q;
  const int ix = (int)(t - q * l.cnt[0]);
  long long iz = (long long)((double)q * l.ry);
  if (iz * l.cnt[1] > q) --iz;
  else if ((iz + 1) * l.cnt[1] <= q) ++iz;
  const int iy = (int)(q - iz * l.cnt[1]);
  const int x = l.start[0] + l.step[0] * ix, y = l.start[1] + l.step[1] * iy, z = l.start[2] + l.step[2] * (int)iz;
  if (l.split) {
    const int key = (x >= a.d[0] - l.r[0] ? 1 : 0) | (y >= a.d[1] - l.r[1] ? 2 : 0);
    if ((key & l.split) != l.key) return;
  }
  const long long m = x + (long long)a.d[0] * (y + (long long)a.d[1] * z);
  const long long nd = a.nd;

  if (OP == SP_VEL2MOM) {
    float u[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) u[i] = (EDGE && a.accum) ? a.u[m + nd * i] : 0.0f;
    // sparse_operator_dev.cu:164-204: channel i sums the lists (i, j) over j in order
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      gfloat vj = a.v + nd * j;
#pragma unroll
      for (int i = 0; i < NC; ++i) u[i] += list_conv<EDGE>(a, i + NC * j, vj, x, y, z, m, a.bnd + 3 * j);
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) a.u[m + nd * i] = u[i];
    return;
  }

  // relax1 / relax_padded1 (sparse_operator_dev.cu:89-161)
  float b[NC], vc[NC];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    b[i] = a.g[m + nd * i];
    vc[i] = a.v[m + nd * i];
  }
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    gfloat vj = a.v + nd * j;
#pragma unroll
    for (int i = 0; i < NC; ++i) b[i] -= list_conv<EDGE>(a, i + NC * j, vj, x, y, z, m, a.bnd + 3 * j);
  }
  float hv[NC * (NC + 1) / 2];
  if (a.hcode) {  // Hv (sparse_operator_dev.cu:32-54)
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      hv[i] = a.h[m + nd * i];
      b[i] -= hv[i] * vc[i];
    }
    if (a.hcode == 2) {
      int ii = NC;
#pragma unroll
      for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int i = j + 1; i < NC; ++i, ++ii) {
          hv[ii] = a.h[m + nd * ii];
          b[j] -= hv[ii] * vc[i];
          b[i] -= hv[ii] * vc[j];
        }
    }
  }
  // make_A (sparse_operator_dev.cu:56-86)
  float A[NC * NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    A[j * NC + j] = a.pairs[j * NC + j].w0 * 1.000001f;
#pragma unroll
    for (int i = j + 1; i < NC; ++i) A[i + j * NC] = A[i * NC + j] = a.pairs[i + j * NC].w0;
  }
  if (a.hcode) {
#pragma unroll
    for (int i = 0; i < NC; ++i) A[i + NC * i] += hv[i] * 1.000001f;
    if (a.hcode == 2) {
      int ii = NC;
#pragma unroll
      for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int i = j + 1; i < NC; ++i, ++ii) A[i + j * NC] = A[j + i * NC] += hv[ii];
    }
  }
  float sol[NC];
  chol_solve<NC>(A, b, sol);
#pragma unroll
  for (int i = 0; i < NC; ++i) a.v[m + nd * i] = vc[i] + sol[i];
}

template <int NC>
hipError_t launch_nc(SpOp op, const SpArgs& a, const SpLattice& l, bool edge, hipStream_t s) {
  const dim3 grid((unsigned)((l.n + kSpBlock - 1) / kSpBlock));
  if (op == SP_VEL2MOM) {
    if (edge)
      hipLaunchKernelGGL((k_sparse<NC, true, SP_VEL2MOM>), grid, dim3(kSpBlock), 0, s, a, l);
    else
      hipLaunchKernelGGL((k_sparse<NC, false, SP_VEL2MOM>), grid, dim3(kSpBlock), 0, s, a, l);
  } else {
    if (edge)
      hipLaunchKernelGGL((k_sparse<NC, true, SP_RELAX>), grid, dim3(kSpBlock), 0, s, a, l);
    else
      hipLaunchKernelGGL((k_sparse<NC, false, SP_RELAX>), grid, dim3(kSpBlock), 0, s, a, l);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_sparse(SpOp op, const SpArgs& a, const SpLattice& l, bool edge, hipStream_t s) {
  if (l.n <= 0) return hipSuccess;
  switch (a.nc) {
    case 1: return launch_nc<1>(op, a, l, edge, s);
    case 2: return launch_nc<2>(op, a, l, edge, s);
    case 3: return launch_nc<3>(op, a, l, edge, s);
    case 4: return launch_nc<4>(op, a, l, edge, s);
    case 5: return launch_nc<5>(op, a, l, edge, s);
    case 6: return launch_nc<6>(op, a, l, edge, s);
    case 7: return launch_nc<7>(op, a, l, edge, s);
    case 8: return launch_nc<8>(op, a, l, edge, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pp
