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
constexpr size_t kSpCoopLds = 48 * 1024;  // LDS budget of the cooperative kernels

// Launches of at most this many voxels run kSpG lanes per voxel: they are too small to
// fill the GPU, and their time is one voxel's latency. Larger ones run a thread per voxel.
constexpr long long kSpCoopMaxVoxels = 32768;

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

// Value of entry t around voxel (x, y, z) (linear index m; mb = 4 m when the volume has
// fewer than 2^30 voxels). Interior voxels read v[m + goff] (BIG = false: one scalar base
// plus a 32-bit byte offset per lane); edge voxels fold each axis by the input channel's
// boundary codes.
template <bool EDGE, bool BIG>
__device__ inline float fetch(const SpArgs& a, const SpEntry& t, int x, int y, int z, long long m, unsigned mb) {
  if (!EDGE) {
    if (BIG) return a.v[m + t.goff];
    return *reinterpret_cast<const PP_GLOBAL float*>(reinterpret_cast<const PP_GLOBAL char*>(a.v + t.goff) + mb);
  }
  const int j = (t.code >> 29) & 7;
  const int dx = (t.code & 255) - 16, dy = ((t.code >> 8) & 255) - 16, dz = ((t.code >> 16) & 255) - 16;
  bool neg = false;
  const int ix = fold_axis(x + dx, a.d[0], a.bnd[3 * j], neg);
  const int iy = fold_axis(y + dy, a.d[1], a.bnd[3 * j + 1], neg);
  const int iz = fold_axis(z + dz, a.d[2], a.bnd[3 * j + 2], neg);
  const float f = a.v[a.nd * j + ix + (long long)a.d[0] * (iy + (long long)a.d[1] * iz)];
  return neg ? -f : f;
}

// All lists of one voxel, in entry order: sp_conv (sparse_operator_dev.cu:12-20) of list
// (i, j) is s = sum_{e>0} w_e (f_e - f0), then s += c0w * f0; it is added to acc[i]
// (vel2mom, sparse_operator_dev.cu:175-181) or subtracted from it (relax, :107-115), for
// j = 0, 1, ... in turn. The taps are fetched kSpBatch at a time (the entry index is
// uniform across the wave, so a batch is kSpBatch independent loads in flight) and then
// accumulated strictly in list order.
constexpr int kSpBatch = 16;

template <int NC, bool EDGE, bool BIG, bool SUB>
__device__ inline void apply_lists(const SpArgs& a, float (&acc)[NC], int x, int y, int z, long long m) {
  const unsigned mb = BIG ? 0u : (unsigned)m * 4u;
  float f0 = 0.0f, s = 0.0f, c0w = 0.0f;
  for (int e0 = 0; e0 < a.nent; e0 += kSpBatch) {
    float f[kSpBatch];
#pragma unroll
    for (int b = 0; b < kSpBatch; ++b) {
      const SpEntry t = a.ent[min(e0 + b, a.nent - 1)];
      f[b] = fetch<EDGE, BIG>(a, t, x, y, z, m, mb);
    }
#pragma unroll
    for (int b = 0; b < kSpBatch; ++b) {
      if (e0 + b < a.nent) {
        const SpEntry t = a.ent[e0 + b];
        if (t.code & kSpFirst) {
          f0 = f[b];
          s = 0.0f;
          c0w = t.w;
        } else {
          s += t.w * (f[b] - f0);
        }
        if (t.code & kSpLast) {
          s += c0w * f0;
          const int i = (t.code >> 26) & 7;
#pragma unroll
          for (int q = 0; q < NC; ++q)
            if (q == i) acc[q] = SUB ? acc[q] - s : acc[q] + s;
        }
      }
    }
  }
}

template <int NC>
__device__ inline void chol_solve(float (&A)[NC * NC], float (&b)[NC], float (&x)[NC]) {
  float p[NC];
  // choldcf (artifacts/C/chol.cu:9-30)
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
  // chollsf (chol.cu:40-59)
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

// H at the voxel (nc or nc (nc + 1) / 2 channels), loaded with g and v before the stencil.
template <int NC>
__device__ inline void load_h(const SpArgs& a, long long m, float (&hv)[NC * (NC + 1) / 2]) {
  if (!a.hcode) return;
  const int nh = a.hcode == 2 ? NC * (NC + 1) / 2 : NC;
#pragma unroll
  for (int i = 0; i < NC * (NC + 1) / 2; ++i)
    if (i < nh) hv[i] = a.h[m + a.nd * i];
}

// relax1 / relax_padded1 after the stencil (sparse_operator_dev.cu:117-131): b -= H v,
// A = the lists' centres (+ H), sol = A \ b by Cholesky. vc: v at the voxel, hv: H there.
template <int NC>
__device__ inline void relax_solve(const SpArgs& a, float (&b)[NC], const float (&vc)[NC],
                                   const float (&hv)[NC * (NC + 1) / 2], float (&sol)[NC]) {
  if (a.hcode) {  // Hv (sparse_operator_dev.cu:32-54)
#pragma unroll
    for (int i = 0; i < NC; ++i) b[i] -= hv[i] * vc[i];
    if (a.hcode == 2) {
      int ii = NC;
#pragma unroll
      for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int i = j + 1; i < NC; ++i, ++ii) {
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
  chol_solve<NC>(A, b, sol);
}

// Lattice index t -> voxel (x, y, z); false when the voxel is outside this launch's phase.
__device__ inline bool lattice_voxel(const SpArgs& a, const SpLattice& l, long long t, int& x, int& y, int& z) {
  // q = t / cnt0 and iz = q / cnt1, exact after one correction
  long long q = (long long)((double)t * l.rx);
  if (q * l.cnt[0] > t) --q;
  else if ((q + 1) * l.cnt[0] <= t) ++q;
  const int ix = (int)(t - q * l.cnt[0]);
  long long iz = (long long)((double)q * l.ry);
  if (iz * l.cnt[1] > q) --iz;
  else if ((iz + 1) * l.cnt[1] <= q) ++iz;
  const int iy = (int)(q - iz * l.cnt[1]);
  x = l.start[0] + l.step[0] * ix;
  y = l.start[1] + l.step[1] * iy;
  z = l.start[2] + l.step[2] * (int)iz;
  if (l.split) {
    const int key = (x >= a.d[0] - l.r[0] ? 1 : 0) | (y >= a.d[1] - l.r[1] ? 2 : 0);
    if ((key & l.split) != l.key) return false;
  }
  return true;
}

// One thread per voxel.
template <int NC, bool EDGE, bool BIG, int OP>
__global__ __launch_bounds__(kSpBlock) void k_sparse(SpArgs a, SpLattice l) {
  const long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * kSpBlock + threadIdx.x;
  if (t >= l.n) return;
  int x, y, z;
  if (!lattice_voxel(a, l, t, x, y, z)) return;
  const long long m = x + (long long)a.d[0] * (y + (long long)a.d[1] * z);
  const long long nd = a.nd;
  if (OP == SP_VEL2MOM) {
    float u[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) u[i] = (EDGE && a.accum) ? a.u[m + nd * i] : 0.0f;
    apply_lists<NC, EDGE, BIG, false>(a, u, x, y, z, m);
#pragma unroll
    for (int i = 0; i < NC; ++i) a.u[m + nd * i] = u[i];
    return;
  }
  float b[NC], vc[NC], sol[NC], hv[NC * (NC + 1) / 2];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    b[i] = a.g[m + nd * i];
    vc[i] = a.v[m + nd * i];
  }
  load_h<NC>(a, m, hv);
  apply_lists<NC, EDGE, BIG, true>(a, b, x, y, z, m);
  relax_solve<NC>(a, b, vc, hv, sol);
#pragma unroll
  for (int i = 0; i < NC; ++i) a.v[m + nd * i] = vc[i] + sol[i];
}

// kSpG lanes per voxel: the lanes fetch the voxel's taps (entry e by lane e mod kSpG) into
// LDS, replace each by its product w_e (f_e - f0), one lane per list then adds its
// products in list order and finishes sp_conv, and lanes 0..NC-1 combine the lists of
// their channel (j in order) and write it. Same operations and order as one thread per
// voxel, with a voxel's memory latency spread over kSpG lanes. The kSpG lanes of a voxel
// lie in one wave.
constexpr int kSpG = 16;
constexpr int kSpVox = kSpBlock / kSpG;
constexpr int kSpLB = 8;  // entries per lane fetched together

template <int NC, bool EDGE, bool BIG, int OP>
__global__ __launch_bounds__(kSpBlock) void k_sparse_coop(SpArgs a, SpLattice l) {
  extern __shared__ float sp_lds[];
  const int g = threadIdx.x / kSpG, r = threadIdx.x % kSpG;
  float* sf = sp_lds + g * a.nent;
  float* cl = sp_lds + kSpVox * a.nent + g * NC * NC;
  const long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * kSpVox + g;
  int x = 0, y = 0, z = 0;
  const bool ok = t < l.n && lattice_voxel(a, l, t, x, y, z);
  const long long m = x + (long long)a.d[0] * (y + (long long)a.d[1] * z);
  const long long nd = a.nd;
  const unsigned mb = BIG ? 0u : (unsigned)m * 4u;
  float b[NC], vc[NC], hv[NC * (NC + 1) / 2];
  if (OP == SP_RELAX && ok && r < NC) {
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      b[i] = a.g[m + nd * i];
      vc[i] = a.v[m + nd * i];
    }
    load_h<NC>(a, m, hv);
  }
  // entries e0 + b kSpG, b < kSpLB, of this lane: all entry loads, then all tap loads
  if (a.nent <= kSpG * kSpLB) {  // one batch per lane: entries and taps stay in registers
    SpEntry en[kSpLB];
    int s0[kSpLB];
    float f[kSpLB];
    if (ok) {
#pragma unroll
      for (int b = 0; b < kSpLB; ++b) {
        en[b] = a.ent[min(r + b * kSpG, a.nent - 1)];
        s0[b] = a.pairs[((en[b].code >> 26) & 7) + NC * ((en[b].code >> 29) & 7)].start;
      }
#pragma unroll
      for (int b = 0; b < kSpLB; ++b) f[b] = fetch<EDGE, BIG>(a, en[b], x, y, z, m, mb);
#pragma unroll
      for (int b = 0; b < kSpLB; ++b)
        if (r + b * kSpG < a.nent) sf[r + b * kSpG] = f[b];
    }
    __syncthreads();
    if (ok) {
#pragma unroll
      for (int b = 0; b < kSpLB; ++b) {
        const int e = r + b * kSpG;
        if (e < a.nent && !(en[b].code & kSpFirst)) sf[e] = en[b].w * (f[b] - sf[s0[b]]);
      }
    }
  } else {
    if (ok)
      for (int e0 = r; e0 < a.nent; e0 += kSpG * kSpLB) {
        SpEntry en[kSpLB];
        float f[kSpLB];
#pragma unroll
        for (int b = 0; b < kSpLB; ++b) en[b] = a.ent[min(e0 + b * kSpG, a.nent - 1)];
#pragma unroll
        for (int b = 0; b < kSpLB; ++b) f[b] = fetch<EDGE, BIG>(a, en[b], x, y, z, m, mb);
#pragma unroll
        for (int b = 0; b < kSpLB; ++b)
          if (e0 + b * kSpG < a.nent) sf[e0 + b * kSpG] = f[b];
      }
    __syncthreads();
    if (ok)
      for (int e0 = r; e0 < a.nent; e0 += kSpG * kSpLB) {
        SpEntry en[kSpLB];
        int s0[kSpLB];
#pragma unroll
        for (int b = 0; b < kSpLB; ++b) {
          en[b] = a.ent[min(e0 + b * kSpG, a.nent - 1)];
          s0[b] = a.pairs[((en[b].code >> 26) & 7) + NC * ((en[b].code >> 29) & 7)].start;
        }
#pragma unroll
        for (int b = 0; b < kSpLB; ++b) {
          const int e = e0 + b * kSpG;
          if (e < a.nent && !(en[b].code & kSpFirst)) sf[e] = en[b].w * (sf[e] - sf[s0[b]]);
        }
      }
  }
  __syncthreads();
  if (ok)
    for (int p = r; p < NC * NC; p += kSpG) {
      const SpPair q = a.pairs[p];
      float s = 0.0f;
      for (int e = q.start + 1; e < q.start + q.len; ++e) s += sf[e];
      s += q.c0w * sf[q.start];
      cl[p] = s;
    }
  __syncthreads();
  if (!ok || r >= NC) return;
  if (OP == SP_VEL2MOM) {
    float u = (EDGE && a.accum) ? a.u[m + nd * r] : 0.0f;
#pragma unroll
    for (int j = 0; j < NC; ++j) u += cl[r + NC * j];
    a.u[m + nd * r] = u;
    return;
  }
#pragma unroll
  for (int j = 0; j < NC; ++j)
#pragma unroll
    for (int i = 0; i < NC; ++i) b[i] -= cl[i + NC * j];
  float sol[NC];
  relax_solve<NC>(a, b, vc, hv, sol);
  float sr = sol[0], vr = vc[0];
#pragma unroll
  for (int i = 1; i < NC; ++i)
    if (r == i) {
      sr = sol[i];
      vr = vc[i];
    }
  a.v[m + nd * r] = vr + sr;
}

// vel2mom over the whole volume, LDS-tiled: a workgroup takes a tile of kVTX x kVTY x kVTZ
// voxels, stages each input channel's tile plus an R-voxel halo in LDS -- every staged
// position folded and signed by that channel's boundary codes, as the reference's
// get_patch does for its edge voxels (and an identity inside) -- and each thread runs kVPT
// voxels (one x, two y, eight z) through the lists of that channel entry by entry, so an
// entry's scalar decoding is shared by kVPT voxels. Per voxel the operations and their
// order are those of the one-thread kernels; with accum, voxels outside the interior block
// [l.start, l.start + l.cnt) start from the u they are given.
constexpr int kVTX = 64, kVTY = 8, kVTZ = 8;
constexpr int kVPT = kVTY * kVTZ / (kSpBlock / kVTX);

template <int NC, int R>
__global__ __launch_bounds__(kSpBlock) void k_vel2mom_tile(SpArgs a, SpLattice l) {
  constexpr int BX = kVTX + 2 * R, BY = kVTY + 2 * R, BZ = kVTZ + 2 * R, NB = BX * BY * BZ;
  __shared__ float box[NB];
  __shared__ int fold[BX + BY + BZ];  // folded coordinate * 2 + sign bit, x then y then z
  const int ntx = (a.d[0] + kVTX - 1) / kVTX, nty = (a.d[1] + kVTY - 1) / kVTY;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
  const int x0 = tx * kVTX, y0 = ty * kVTY, z0 = tz * kVTZ;
  const int lx = threadIdx.x % kVTX, lg = threadIdx.x / kVTX;  // lg: y rows 2 lg, 2 lg + 1
  const long long nd = a.nd;
  const int x = x0 + lx;
  float u[kVPT][NC];
#pragma unroll
  for (int v = 0; v < kVPT; ++v) {
    const int y = y0 + 2 * lg + v % 2, z = z0 + v / 2;
    const bool edge = x < l.start[0] || x >= l.start[0] + l.cnt[0] || y < l.start[1] || y >= l.start[1] + l.cnt[1] ||
                      z < l.start[2] || z >= l.start[2] + l.cnt[2];
    const bool in = x < a.d[0] && y < a.d[1] && z < a.d[2];
    const long long m = x + (long long)a.d[0] * (y + (long long)a.d[1] * z);
#pragma unroll
    for (int i = 0; i < NC; ++i) u[v][i] = (a.accum && edge && in) ? a.u[m + nd * i] : 0.0f;
  }
  const int base = lx + BX * 2 * lg;  // box corner of voxel 0 of this thread
  // the box lies inside the volume: every fold is the identity
  const bool inside = x0 >= R && y0 >= R && z0 >= R && x0 + BX - R <= a.d[0] && y0 + BY - R <= a.d[1] &&
                      z0 + BZ - R <= a.d[2];
  constexpr int kStage = 16, kIter = (NB + kSpBlock - 1) / kSpBlock;
  for (int j = 0; j < NC; ++j) {
    if (j) __syncthreads();
    const PP_GLOBAL float* vj = a.v + nd * j;
    if (!inside) {  // fold tables of this channel's codes for the box's x, y and z ranges
      for (int t = threadIdx.x; t < BX + BY + BZ; t += kSpBlock) {
        const int ax = t < BX ? 0 : (t < BX + BY ? 1 : 2);
        const int c = ax == 0 ? x0 - R + t : (ax == 1 ? y0 - R + (t - BX) : z0 - R + (t - BX - BY));
        bool neg = false;
        const int g = fold_axis(c, a.d[ax], a.bnd[3 * j + ax], neg);
        fold[t] = 2 * g + (neg ? 1 : 0);
      }
      __syncthreads();
    }
    // staging: kStage loads in flight per thread, then their LDS stores
    if (inside) {
      const PP_GLOBAL float* o = vj + (x0 - R) + (long long)a.d[0] * ((y0 - R) + (long long)a.d[1] * (z0 - R));
#pragma unroll 1
      for (int it = 0; it < kIter; it += kStage) {
        float t[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
          const int e = min((it + q) * kSpBlock + (int)threadIdx.x, NB - 1);
          const int bx = e % BX, by = (e / BX) % BY, bz = e / (BX * BY);
          t[q] = o[bx + (long long)a.d[0] * (by + (long long)a.d[1] * bz)];
        }
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
          const int e = (it + q) * kSpBlock + (int)threadIdx.x;
          if (it + q < kIter && e < NB) box[e] = t[q];
        }
      }
    } else {
#pragma unroll 1
      for (int it = 0; it < kIter; it += kStage) {
        float t[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
          const int e = min((it + q) * kSpBlock + (int)threadIdx.x, NB - 1);
          const int bx = e % BX, by = (e / BX) % BY, bz = e / (BX * BY);
          const int fx = fold[bx], fy = fold[BX + by], fz = fold[BX + BY + bz];
          const float f = vj[(fx >> 1) + (long long)a.d[0] * ((fy >> 1) + (long long)a.d[1] * (fz >> 1))];
          t[q] = ((fx ^ fy ^ fz) & 1) ? -f : f;
        }
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
          const int e = (it + q) * kSpBlock + (int)threadIdx.x;
          if (it + q < kIter && e < NB) box[e] = t[q];
        }
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int i = 0; i < NC; ++i) {
      const SpPair q = a.pairs[i + NC * j];
      float f0[kVPT], sv[kVPT];
      {
        const int c = a.ent[q.start].code;
        const int o = ((c & 255) - 16 + R) + BX * ((((c >> 8) & 255) - 16 + R) + BY * (((c >> 16) & 255) - 16 + R));
#pragma unroll
        for (int v = 0; v < kVPT; ++v) {
          f0[v] = box[base + o + BX * ((v % 2) + BY * (v / 2))];
          sv[v] = 0.0f;
        }
      }
#pragma unroll 2
      for (int e = q.start + 1; e < q.start + q.len; ++e) {
        const SpEntry t = a.ent[e];
        const int c = t.code;
        const int o = ((c & 255) - 16 + R) + BX * ((((c >> 8) & 255) - 16 + R) + BY * (((c >> 16) & 255) - 16 + R));
#pragma unroll
        for (int v = 0; v < kVPT; ++v) sv[v] += t.w * (box[base + o + BX * ((v % 2) + BY * (v / 2))] - f0[v]);
      }
#pragma unroll
      for (int v = 0; v < kVPT; ++v) {
        sv[v] += q.c0w * f0[v];
#pragma unroll
        for (int ii = 0; ii < NC; ++ii)
          if (ii == i) u[v][ii] += sv[v];
      }
    }
  }
  if (x >= a.d[0]) return;
#pragma unroll
  for (int v = 0; v < kVPT; ++v) {
    const int y = y0 + 2 * lg + v % 2, z = z0 + v / 2;
    if (y < a.d[1] && z < a.d[2]) {
      const long long m = x + (long long)a.d[0] * (y + (long long)a.d[1] * z);
#pragma unroll
      for (int i = 0; i < NC; ++i) a.u[m + nd * i] = u[v][i];
    }
  }
}

template <int NC>
void launch_vtile(const SpArgs& a, const SpLattice& l, int r, hipStream_t s) {
  const long long nt = (long long)((a.d[0] + kVTX - 1) / kVTX) * ((a.d[1] + kVTY - 1) / kVTY) *
                       ((a.d[2] + kVTZ - 1) / kVTZ);
  if (r == 1)
    hipLaunchKernelGGL((k_vel2mom_tile<NC, 1>), dim3((unsigned)nt), dim3(kSpBlock), 0, s, a, l);
  else
    hipLaunchKernelGGL((k_vel2mom_tile<NC, 2>), dim3((unsigned)nt), dim3(kSpBlock), 0, s, a, l);
}

template <int NC, int OP>
void launch_op(const SpArgs& a, const SpLattice& l, bool edge, hipStream_t s) {
  const size_t lds = (size_t)kSpVox * (a.nent + NC * NC) * sizeof(float);
  if (lds <= kSpCoopLds && l.n <= kSpCoopMaxVoxels) {
    const dim3 grid((unsigned)((l.n + kSpVox - 1) / kSpVox));
    if (edge)
      hipLaunchKernelGGL((k_sparse_coop<NC, true, true, OP>), grid, dim3(kSpBlock), lds, s, a, l);
    else if (a.nd >= (1LL << 30))
      hipLaunchKernelGGL((k_sparse_coop<NC, false, true, OP>), grid, dim3(kSpBlock), lds, s, a, l);
    else
      hipLaunchKernelGGL((k_sparse_coop<NC, false, false, OP>), grid, dim3(kSpBlock), lds, s, a, l);
    return;
  }
  const dim3 grid((unsigned)((l.n + kSpBlock - 1) / kSpBlock));
  if (edge)
    hipLaunchKernelGGL((k_sparse<NC, true, true, OP>), grid, dim3(kSpBlock), 0, s, a, l);
  else if (a.nd >= (1LL << 30))
    hipLaunchKernelGGL((k_sparse<NC, false, true, OP>), grid, dim3(kSpBlock), 0, s, a, l);
  else
    hipLaunchKernelGGL((k_sparse<NC, false, false, OP>), grid, dim3(kSpBlock), 0, s, a, l);
}

template <int NC>
hipError_t launch_nc(SpOp op, const SpArgs& a, const SpLattice& l, bool edge, hipStream_t s) {
  if (op == SP_VEL2MOM)
    launch_op<NC, SP_VEL2MOM>(a, l, edge, s);
  else
    launch_op<NC, SP_RELAX>(a, l, edge, s);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_vel2mom_tiled(const SpArgs& a, const SpLattice& interior, int r, hipStream_t s) {
  if (r < 1 || r > 2) return hipErrorNotSupported;
  const SpLattice& l = interior;
  switch (a.nc) {
    case 1: launch_vtile<1>(a, l, r, s); break;
    case 2: launch_vtile<2>(a, l, r, s); break;
    case 3: launch_vtile<3>(a, l, r, s); break;
    default: return hipErrorNotSupported;
  }
  return hipGetLastError();
}

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
