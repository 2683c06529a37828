// pp_tiled.hip -- tiled gfx950 kernels for the hot ops (pull, push, push_grad, the phi
// pullback of pull, and the fused matching objective).
//
// A 256-thread workgroup owns a TX x TY x TZ tile of the phi grid (TX*TY*TZ/256 samples
// per thread). Its samples' knots are reduced to a bounding box in the *unfolded* index
// space of the image grid; the box (<= CAP cells) lives in LDS:
//   gathers  stage the box once (coalesced row loads, boundary fold and sign applied
//            while staging), then every sample reads its dp^3 taps from LDS;
//   scatters accumulate every tap in fixed point with an LDS integer atomic (32-bit
//            cells for values, 64-bit for push_grad), then flush each non-zero cell
//            with one folded global atomic -- a few global atomics per sample instead
//            of dp^3, and none of the same-address contention of neighbouring samples
//            hitting the same HBM lines.
// Samples whose knot lies outside a window around the tile's centre sample (displaced
// far away, sent outside the field of view, or on the large-coordinate path) are
// handled individually through global memory, so a few outliers never inflate the box.
// Tiles whose box would exceed CAP fall back to the global path entirely.
#define PP_KERNEL_SOURCE 1  // kernel arguments carry the global address space (pp_kernels.hpp)
#include "pp_sample.hpp"

#include <algorithm>
#include <climits>

namespace pp {

constexpr int kTB = 256;

// Largest box extent along y or z: 2 * (window half-width) + dp for every tile shape.
constexpr int kMaxYZ = 64;

struct TileState {
  int lo[3];      // unfolded image index of box cell (0,0,0), per axis
  int b[3];       // box extent per axis
  int cells;      // b[0]*b[1]*b[2]
  int lds;        // 1 when the box is used
  int ctr[3];     // window centre (knot of the tile's centre sample)
  int org[3];     // tile origin in the phi grid
  int red[6][4];  // per-wave partial min/max
  float redf[4];  // per-wave partial bound (tile_setup's `bound` hook)
  // folded image offset of box plane y = i (d0x * fold(lo[1] + i)) and z = i
  // (d0x * d0y * fold(lo[2] + i)); bit 31 set where the boundary flips the sign
  int yoff[kMaxYZ], zoff[kMaxYZ];
};

template <int TX, int TY, int TZ>
struct Shape {
  static constexpr int N = TX * TY * TZ;
  static constexpr int VPT = N / kTB;
  static_assert(N % kTB == 0, "a tile is a whole number of workgroup passes");
  // window half-width: a knot farther than this from the centre sample's is an outlier
  static constexpr int WX = TX + 8, WY = TY + 8, WZ = TZ + 8;
  // a box row (<= 2 * WX + dp cells) fits in one workgroup pass (box_walk relies on it)
  static_assert(2 * WX + 5 < kTB, "box rows must be shorter than the workgroup");
  static_assert(2 * WY + 5 <= kMaxYZ && 2 * WZ + 5 <= kMaxYZ, "box planes fit the fold tables");
  // local index of the centre sample
  static constexpr int LC = TX / 2 + TX * (TY / 2 + TY * (TZ / 2));
};

// Cyclic rotation by one element. Loops over a thread's samples that are not unrolled
// read element 0 and rotate once per iteration (VPT rotations restore the order), so the
// per-sample arrays stay in registers: a runtime index would put them in scratch memory.
template <typename T, int N>
__device__ __forceinline__ void rot(T (&a)[N]) {
  const T t = a[0];
#pragma unroll
  for (int i = 0; i < N - 1; ++i) a[i] = a[i + 1];
  a[N - 1] = t;
}

template <typename T, int M, int N>
__device__ __forceinline__ void rot_rows(T (&a)[M][N]) {
#pragma unroll
  for (int m = 0; m < M; ++m) rot(a[m]);
}

template <int VPT>
struct Vox {
  float x[VPT], y[VPT], z[VPT];
  int v[VPT];
  unsigned inside, valid, inl;  // bit q: position inside the phi grid / valid sample / in-window
  __device__ __forceinline__ void rotate() {
    rot(x);
    rot(y);
    rot(z);
    rot(v);
  }
};

// Wave-wide reductions (DPP lane moves, not LDS-routed permutes).
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }

template <int DP>
__device__ __forceinline__ int knot_fast(float x) {
  return (int)ceilf(x - 0.5f * (float)DP);
}

__device__ __forceinline__ bool is_fast3(float x, float y, float z) {
  return fabsf(x) < kFastLimit && fabsf(y) < kFastLimit && fabsf(z) < kFastLimit;
}

typedef unsigned long long acc_t;  // fixed-point LDS accumulator cell (scatters)

// No-op hooks of tile_setup.
struct NoPrefetch {
  template <class VoxT>
  __device__ __forceinline__ void operator()(const VoxT&) const {}
};
struct NoBound {
  __device__ __forceinline__ float operator()() const { return 0.f; }
};

// Load the tile's samples, classify them and size the LDS box. Ends with a barrier.
// Hooks, so that callers share its two barriers instead of adding their own:
//   pre(V)    runs once the samples' grid indices are known, before the deformation is
//             read: loads issued there overlap the deformation's;
//   bound()   a per-thread value >= 0 whose workgroup maximum (NARROW: sum) lands in
//             *block_bound;
//   zero_box  when given, the box's cells (NARROW: 32-bit, else 64-bit) are zeroed
//             before the final barrier.
template <int DP, int TX, int TY, int TZ, int CAP, bool NARROW = false, class Pre = NoPrefetch,
          class Bound = NoBound>
__device__ __forceinline__ void tile_setup(const KArgs& a, const float* __restrict__ phi, int tile,
                                           Vox<Shape<TX, TY, TZ>::VPT>& V, TileState& S, Pre pre = Pre(),
                                           Bound bound = Bound(), acc_t* zero_box = nullptr,
                                           float* block_bound = nullptr) {
  using SH = Shape<TX, TY, TZ>;
  constexpr int VPT = SH::VPT;
  const int ntx = (a.d1x + TX - 1) / TX, nty = (a.d1y + TY - 1) / TY;
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int tid = threadIdx.x;
  V.inside = 0;
  V.valid = 0;
  V.inl = 0;
#pragma unroll
  for (int q = 0; q < VPT; ++q) {
    const int l = tid + kTB * q;
    const int gx = x0 + l % TX, gy = y0 + (l / TX) % TY, gz = z0 + l / (TX * TY);
    const bool in = gx < a.d1x && gy < a.d1y && gz < a.d1z;
    V.v[q] = in ? gx + a.d1x * (gy + a.d1y * gz) : 0;
    if (in) V.inside |= 1u << q;
  }
  pre(V);
#pragma unroll
  for (int q = 0; q < VPT; ++q) {
    const int l = tid + kTB * q;
    const int gx = x0 + l % TX, gy = y0 + (l / TX) % TY, gz = z0 + l / (TX * TY);
    float x = 0.f, y = 0.f, z = 0.f;
    bool ok = false;
    if ((V.inside >> q) & 1u) {
      sample_coords_at(a, phi, V.v[q], gx, gy, gz, x, y, z);
      ok = voxok(x, y, z, a.ext, a.lx, a.ly, a.lz);
    }
    V.x[q] = x;
    V.y[q] = y;
    V.z[q] = z;
    if (ok) V.valid |= 1u << q;
    if (l == SH::LC) {  // window centre: the centre sample's knot, else the tile centre
      S.org[0] = x0;    // mapped by the grid-size ratio
      S.org[1] = y0;
      S.org[2] = z0;
      if (ok && is_fast3(x, y, z)) {
        S.ctr[0] = knot_fast<DP>(x);
        S.ctr[1] = knot_fast<DP>(y);
        S.ctr[2] = knot_fast<DP>(z);
      } else {  // (double: exact for grids below 2^26 per axis, and no 64-bit division)
        S.ctr[0] = (int)((double)(x0 + TX / 2) * a.d0x / a.d1x) - (DP - 1) / 2;
        S.ctr[1] = (int)((double)(y0 + TY / 2) * a.d0y / a.d1y) - (DP - 1) / 2;
        S.ctr[2] = (int)((double)(z0 + TZ / 2) * a.d0z / a.d1z) - (DP - 1) / 2;
      }
    }
  }
  __syncthreads();
  const int cx = S.ctr[0], cy = S.ctr[1], cz = S.ctr[2];
  int mn0 = INT_MAX, mn1 = INT_MAX, mn2 = INT_MAX, mx0 = INT_MIN, mx1 = INT_MIN, mx2 = INT_MIN;
#pragma unroll
  for (int q = 0; q < VPT; ++q) {
    if (!((V.valid >> q) & 1u) || !is_fast3(V.x[q], V.y[q], V.z[q])) continue;
    const int ox = knot_fast<DP>(V.x[q]), oy = knot_fast<DP>(V.y[q]), oz = knot_fast<DP>(V.z[q]);
    if (abs(ox - cx) > SH::WX || abs(oy - cy) > SH::WY || abs(oz - cz) > SH::WZ) continue;
    V.inl |= 1u << q;
    mn0 = min(mn0, ox); mx0 = max(mx0, ox);
    mn1 = min(mn1, oy); mx1 = max(mx1, oy);
    mn2 = min(mn2, oz); mx2 = max(mx2, oz);
  }
  mn0 = wave_min(mn0); mn1 = wave_min(mn1); mn2 = wave_min(mn2);
  mx0 = wave_max(mx0); mx1 = wave_max(mx1); mx2 = wave_max(mx2);
  float bnd = bound();
  if (block_bound)
    bnd = NARROW ? __ockl_wfred_add_f32(bnd) : __ockl_wfred_max_f32(bnd);
  const int wid = tid >> 6;
  if ((tid & 63) == 0) {
    S.red[0][wid] = mn0; S.red[1][wid] = mn1; S.red[2][wid] = mn2;
    S.red[3][wid] = mx0; S.red[4][wid] = mx1; S.red[5][wid] = mx2;
    S.redf[wid] = bnd;
  }
  __syncthreads();
  if (block_bound) {
    float m = S.redf[0];
#pragma unroll
    for (int w = 1; w < kTB / 64; ++w) m = NARROW ? m + S.redf[w] : fmaxf(m, S.redf[w]);
    *block_bound = m;
  }
  // every thread sizes the box (a broadcast read of the partials)
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
  for (int w = 0; w < kTB / 64; ++w)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      lo[q] = min(lo[q], S.red[q][w]);
      hi[q] = max(hi[q], S.red[3 + q][w]);
    }
  const bool any = lo[0] <= hi[0];  // some in-window sample
  const int by = hi[1] - lo[1] + DP, bz = hi[2] - lo[2] + DP;
  const int cells = any ? (hi[0] - lo[0] + DP) * by * bz : 0;
  const bool use = any && cells <= CAP;
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      S.lo[q] = lo[q];
      S.b[q] = hi[q] - lo[q] + DP;
    }
    S.cells = cells;
    S.lds = use;
  }
  if (use && zero_box)  // (32-bit cells: as many 64-bit words as cover them)
    for (int i = tid; i < (NARROW ? (cells + 1) >> 1 : cells); i += kTB) zero_box[i] = 0ull;
  if (use) {  // fold tables of the box planes
    bool neg;
    if (tid < by) {
      const int f = fold32(a.by, lo[1] + tid, a.d0y, neg);
      S.yoff[tid] = (a.d0x * f) | (neg ? INT_MIN : 0);
    } else if (tid >= 128 && tid - 128 < bz) {
      const int f = fold32(a.bz, lo[2] + tid - 128, a.d0z, neg);
      S.zoff[tid - 128] = (a.d0x * a.d0y * f) | (neg ? INT_MIN : 0);
    }
  }
  __syncthreads();
}

// The box is walked by the whole workgroup: thread t owns column t % BX of rows
// t / BX, t / BX + RP, ... (RP = kTB / BX rows per pass; BX <= 2 * WX + DP < kTB; rows
// run over y, then z). Row coordinates advance incrementally and the folds of y and z
// come from the tile's tables, so a cell costs a few instructions. kBatch rows' loads
// are issued before any is used, so a wave has that many in flight.
constexpr int kBatch = 16;

struct BoxWalk {
  int c, rp;       // column, rows per pass
  int ry, rz;      // plane indices of the current row
  int dy, dz;      // plane step per pass: RP = dz * BY + dy
  int r, rows;     // current row, rows in the box
  int fx;          // folded x offset of the column
  bool nx;         // sign flip of the column
  bool active;     // r0 < RP: this thread owns a column
};

__device__ __forceinline__ BoxWalk box_walk(const TileState& S, const KArgs& a) {
  const int BX = S.b[0], BY = S.b[1];
  BoxWalk w;
  w.c = threadIdx.x % BX;
  const int r0 = threadIdx.x / BX;
  w.rp = kTB / BX;
  w.active = r0 < w.rp;
  w.r = r0;
  w.rows = BY * S.b[2];
  w.ry = r0 % BY;
  w.rz = r0 / BY;
  w.dy = w.rp % BY;
  w.dz = w.rp / BY;
  w.fx = fold32(a.bx, S.lo[0] + w.c, a.d0x, w.nx);
  return w;
}

__device__ __forceinline__ void walk_next(BoxWalk& w, int BY) {
  w.r += w.rp;
  w.ry += w.dy;
  w.rz += w.dz;
  if (w.ry >= BY) {
    w.ry -= BY;
    ++w.rz;
  }
}

// Folded image offset of the walk's current cell and whether its sign flips; `signs`
// (bit a) selects the axes whose boundary sign applies to the channel.
__device__ __forceinline__ int walk_offset(const TileState& S, const BoxWalk& w, unsigned signs, bool& neg) {
  const int yo = S.yoff[w.ry], zo = S.zoff[w.rz];
  neg = ((yo < 0) && (signs & 2u)) != ((zo < 0) && (signs & 4u));
  neg = neg != (w.nx && (signs & 1u));
  return w.fx + (yo & INT_MAX) + (zo & INT_MAX);
}

// Stage the box from image channel `src` (fold + boundary sign applied here).
__device__ __forceinline__ void stage_box(float* box, const TileState& S, const float* __restrict__ src,
                                          const KArgs& a, unsigned signs) {
  BoxWalk w = box_walk(S, a);
  if (!w.active) return;
  const int BX = S.b[0], BY = S.b[1];
  while (w.r < w.rows) {
    const int r0 = w.r;
    float v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      v[u] = 0.f;
      if (w.r < w.rows) {
        bool neg;
        const float t = src[walk_offset(S, w, signs, neg)];
        v[u] = neg ? -t : t;
      }
      walk_next(w, BY);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int r = r0 + u * w.rp;
      if (r < w.rows) box[r * BX + w.c] = v[u];
    }
  }
}

// ---- fixed-point scatter accumulation ---------------------------------------------
// gfx950 executes float atomics (LDS ds_add_f32 and global alike) at a small fraction
// of the rate of integer LDS atomics, so scatters accumulate into a fixed-point LDS
// box with integer adds: every contribution c is stored as round(c * 2^scale). In
// 64-bit cells (ds_add_u64), per channel and tile, scale = 50 - e where 2^e bounds the
// largest contribution: every scaled contribution is below 2^50 and no cell can
// overflow (at most 2^10 samples per tile touch a cell). Integer sums are exact and
// order-independent, and each cell is rounded to float once, when it is flushed.

// A channel's conversion factor mul: contribution c -> round(c * mul); cell q -> q / mul.
__device__ __forceinline__ double fixed_scale(float vmax) {
  int e = 50;                               // (vmax = 0: scale 0)
  if (vmax > 0.0f) (void)frexpf(vmax, &e);  // vmax < 2^e
  return ldexp(1.0, 50 - e);                // a power of two: products stay exact
}

// Scatters of values (push, and the residual scatter of the objective and of the VJP)
// use 32-bit cells and ds_add_u32, which retires ~1.3x as many lanes per clock as
// ds_add_u64. Wrapping two's-complement sums are exact in any order, so only a cell's
// final value has to fit: no cell exceeds B = wmax^3 * sum|v| over the tile's samples
// (wmax the largest 1-D B-spline weight, every tap weight lies in [0, wmax^3]). With
// mul = (2^31 - 2^10) / B a cell stays below 2^31 even with the rounding of its at most
// 2^10 contributions added. mul is no power of two: the scaled products round once more
// in double (2^-53 relative, far below the quantum). The quantum 1/mul is at most
// 2^-31 * wmax^3 * 1024 times the tile's largest value (1.4e-7 of it for cubic splines)
// and finer where values vary. push_grad keeps 64-bit cells.
template <bool GRAD>
constexpr bool narrow_cells() {
  return !GRAD;
}

template <int DP>
constexpr float wmax3() {  // (largest weight of the degree DP-1 B-spline)^3
  return DP <= 2 ? 1.0f : DP == 3 ? 0.421875f : DP == 4 ? 0.2962963f : 0.2148771f;
}

template <int DP>
__device__ __forceinline__ double fixed_scale32(float vsum) {
  // (1.0001: the float sum and the float weights may fall short of the exact bound by
  // a few parts in 10^5; a sum beyond the float range is below 2^10 * FLT_MAX)
  const double s = vsum <= __FLT_MAX__ ? (double)vsum : 1024.0 * (double)__FLT_MAX__;
  const double bound = s * (double)wmax3<DP>() * 1.0001;
  return bound > 0.0 ? (2147483648.0 - 1024.0) / bound : 1.0;
}

template <bool NARROW>
struct CellOf {
  typedef acc_t T;
};
template <>
struct CellOf<true> {
  typedef unsigned T;
};

template <class T>
struct Cell;
template <>
struct Cell<acc_t> {
  static __device__ __forceinline__ long long q(acc_t v) { return (long long)v; }
};
template <>
struct Cell<unsigned> {
  static __device__ __forceinline__ long long q(unsigned v) { return (long long)(int)v; }
};

// round(a * b) as int64 for |a * b| < 2^51: the FMA with 2^52 + 2^51 rounds the product
// to an integer held in the low mantissa bits (one FMA instead of a float -> int64
// conversion sequence); fix2 adds two such products.
constexpr double kMagic = 6755399441055744.0;  // 2^52 + 2^51

__device__ __forceinline__ acc_t fix1(double a, double b) {
  return (acc_t)(__double_as_longlong(fma(a, b, kMagic)) - __double_as_longlong(kMagic));
}
__device__ __forceinline__ acc_t fix2(double a, double b, double c, double d) {
  return (acc_t)(__double_as_longlong(fma(a, b, fma(c, d, kMagic))) - __double_as_longlong(kMagic));
}

// q * inv rounded to float: below 2^51 in magnitude q converts exactly through the same
// magic constant (as fix1, in reverse); larger sums take the general conversion.
__device__ __forceinline__ float fixed_to_float(long long q, double inv) {
  if (q > -(1LL << 51) && q < (1LL << 51)) {
    const double d = __longlong_as_double(q + __double_as_longlong(kMagic)) - kMagic;
    return (float)(d * inv);
  }
  return (float)((double)q * inv);
}

template <class T>
__device__ __forceinline__ void zero_acc(T* box, int cells) {
  for (int i = threadIdx.x; i < cells; i += kTB) box[i] = 0;
}

// Largest value (SUM: the sum) over the workgroup (v >= 0); uses and releases wsum.
template <bool SUM = false>
__device__ __forceinline__ float block_max(float v, float* wsum) {
  v = SUM ? __ockl_wfred_add_f32(v) : __ockl_wfred_max_f32(v);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  float m = wsum[0];
  for (int w = 1; w < kTB / 64; ++w) m = SUM ? m + wsum[w] : fmaxf(m, wsum[w]);
  __syncthreads();
  return m;
}

// Flush the fixed-point box into image channel `dst`: one folded global float atomic per
// non-zero cell.
template <class T>
__device__ __forceinline__ void flush_acc(const T* box, const TileState& S, double mul, float* __restrict__ dst,
                                          const KArgs& a, unsigned signs) {
  BoxWalk w = box_walk(S, a);
  if (!w.active) return;
  const int BX = S.b[0], BY = S.b[1];
  const double inv = 1.0 / mul;  // (exact for the powers of two of 64-bit cells)
  for (; w.r < w.rows; walk_next(w, BY)) {
    const long long q = Cell<T>::q(box[w.r * BX + w.c]);
    if (q != 0) {
      bool neg;
      const int off = walk_offset(S, w, signs, neg);
      const float v = fixed_to_float(q, inv);
      unsafeAtomicAdd(dst + off, neg ? -v : v);
    }
  }
}

// Unsigned, unfolded weights of one in-window sample and its box offset
template <int DP>
struct LSample {
  float wx[DP], wy[DP], wz[DP], dx[DP], dy[DP], dz[DP];
  int base;
};

template <int DP, bool DERIV>
__device__ __forceinline__ void lsample(float x, float y, float z, const TileState& S, LSample<DP>& L) {
  const int ox = knot_fast<DP>(x), oy = knot_fast<DP>(y), oz = knot_fast<DP>(z);
  float h[DP];
  weights_ct<DP>(x - (float)ox, L.wx, L.dx, h);
  weights_ct<DP>(y - (float)oy, L.wy, L.dy, h);
  weights_ct<DP>(z - (float)oz, L.wz, L.dz, h);
  L.base = (ox - S.lo[0]) + S.b[0] * ((oy - S.lo[1]) + S.b[1] * (oz - S.lo[2]));
}

// ... with the second-derivative weights too (pull_hess and its pullback)
template <int DP>
struct LSampleH : LSample<DP> {
  float hx[DP], hy[DP], hz[DP];
};

template <int DP>
__device__ __forceinline__ void lsample_h(float x, float y, float z, const TileState& S, LSampleH<DP>& L) {
  const int ox = knot_fast<DP>(x), oy = knot_fast<DP>(y), oz = knot_fast<DP>(z);
  weights_ct<DP>(x - (float)ox, L.wx, L.dx, L.hx);
  weights_ct<DP>(y - (float)oy, L.wy, L.dy, L.hy);
  weights_ct<DP>(z - (float)oz, L.wz, L.dz, L.hz);
  L.base = (ox - S.lo[0]) + S.b[0] * ((oy - S.lo[1]) + S.b[1] * (oz - S.lo[2]));
}

// ---- push / push_grad -------------------------------------------------------------
// One tile of push (GRAD=false) or push_grad (GRAD=true) for batch index bb. The caller
// provides the LDS box (CAP cells: 32-bit for push, 64-bit for push_grad, 8-byte aligned),
// wsum (kTB/64 floats) and the tile state.
template <int DP, int TX, int TY, int TZ, int CAP, bool GRAD>
__device__ __forceinline__ void push_tile(const KArgs& a, int tile, int bb, acc_t* box, float* wsum, TileState& S) {
  constexpr int VPT = Shape<TX, TY, TZ>::VPT;
  constexpr int NG = GRAD ? 3 : 1;
  constexpr bool NW = narrow_cells<GRAD>();
  typedef typename CellOf<NW>::T C;
  C* cbox = reinterpret_cast<C*>(box);
  const float* phi = a.phi + bb * a.phi_bs;
  const long long n1 = a.n1;
  // the first channel's values are loaded before the deformation is read
  float g[NG][VPT];
  const float* src0 = a.f + bb * a.f_bs;
  const bool read0 = GRAD || (a.vmode == 0 && a.nchan > 0);
  auto prefetch = [&](const Vox<VPT>& W) {
#pragma unroll
    for (int q = 0; q < VPT; ++q)
#pragma unroll
      for (int e = 0; e < NG; ++e) g[e][q] = (read0 && ((W.inside >> q) & 1u)) ? src0[W.v[q] + e * n1] : 0.f;
  };
  // channel 0's scale bound: the largest (32-bit cells: the sum of the) finite per-sample
  // bound over all samples in the grid (a superset of the in-window ones, so a valid
  // bound), reduced with the box
  auto bound0 = [&]() {
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      float b = fabsf(g[0][q]);
      if (GRAD) b += fabsf(g[NG > 1 ? 1 : 0][q]) + fabsf(g[NG > 2 ? 2 : 0][q]);
      if (isfinite(b)) m = NW ? m + b : fmaxf(m, b);
    }
    return m;
  };
  Vox<VPT> V;
  float bound_c0 = 0.f;
  tile_setup<DP, TX, TY, TZ, CAP, NW>(a, phi, tile, V, S, prefetch, bound0, box, &bound_c0);
  const bool lds = S.lds != 0;
  const int BX = S.b[0], BXY = S.b[0] * S.b[1];
  // push with a count image (out2): one more channel of ones after the nchan channels
  const int nsrc = a.nchan + ((!GRAD && a.out2) ? 1 : 0);
  for (int c = 0; c < nsrc; ++c) {
    const bool count = c == a.nchan;
    const int vmode = GRAD ? 0 : count ? 1 : a.vmode;  // push_grad reads its gradients
    const float* src = a.f + bb * a.f_bs + c * a.f_cs;
    float* dst = count ? a.out2 + bb * a.out2_bs : a.out + bb * a.out_bs + c * a.out_cs;
    // this channel's values of all the thread's valid samples, loaded together
    if (!(c == 0 && read0)) {
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        const int l = threadIdx.x + kTB * q;
        const int ax = c + a.chan0;
        const int id = ax == 0 ? S.org[0] + l % TX : ax == 1 ? S.org[1] + (l / TX) % TY : S.org[2] + l / (TX * TY);
#pragma unroll
        for (int e = 0; e < NG; ++e)
          g[e][q] = !((V.valid >> q) & 1u) ? 0.f
                    : vmode == 0           ? src[V.v[q] + e * n1]
                    : vmode == 1           ? 1.0f
                                           : (float)(id + 1);
      }
    }
    unsigned use = lds ? V.inl : 0u;  // samples accumulated in LDS for this channel
    double scale = 1.0;
    if (lds) {
      // bound on |contribution| (|w|, |dw| <= 1); non-finite values take the global path
      float vmax = 0.f;
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        float b = fabsf(g[0][q]);
        if (GRAD) b += fabsf(g[NG > 1 ? 1 : 0][q]) + fabsf(g[NG > 2 ? 2 : 0][q]);
        if (!((use >> q) & 1u)) continue;
        if (isfinite(b)) vmax = NW ? vmax + b : fmaxf(vmax, b);
        else use &= ~(1u << q);
      }
      if (c == 0 && read0) {  // bound reduced and box zeroed within tile_setup's barriers
        scale = NW ? fixed_scale32<DP>(bound_c0) : fixed_scale(bound_c0);
      } else {
        const float m = block_max<NW>(vmax, wsum);
        scale = NW ? fixed_scale32<DP>(m) : fixed_scale(m);
        zero_acc(cbox, S.cells);
        __syncthreads();
      }
    }
#pragma unroll 1
    for (int q = 0; q < VPT; ++q, V.rotate(), rot_rows(g)) {
      if (!((V.valid >> q) & 1u)) continue;
      const float g0 = g[0][0];
      const float g1 = GRAD ? g[NG > 1 ? 1 : 0][0] : 0.f;
      const float g2 = GRAD ? g[NG > 2 ? 2 : 0][0] : 0.f;
      if (!GRAD && !isfinite(g0)) continue;  // push skips non-finite values (pushpull.c:155)
      const float x = V.x[0], y = V.y[0], z = V.z[0];
      if ((use >> q) & 1u) {
        LSample<DP> L;
        lsample<DP, GRAD>(x, y, z, S, L);
        const double s0 = (double)g0 * scale;
        const double s1 = GRAD ? (double)g1 * scale : 0.0;
        const double s2 = GRAD ? (double)g2 * scale : 0.0;
#pragma unroll
        for (int k = 0; k < DP; ++k)
#pragma unroll
          for (int j = 0; j < DP; ++j) {
            C* row = cbox + L.base + BX * j + BXY * k;
            if (GRAD) {
              const double cx = (double)(L.wz[k] * L.wy[j]) * s0;
              const double cyz = (double)(L.wz[k] * L.dy[j]) * s1 + (double)(L.dz[k] * L.wy[j]) * s2;
#pragma unroll
              for (int i = 0; i < DP; ++i) atomicAdd(row + i, (C)fix2(L.dx[i], cx, L.wx[i], cyz));
            } else {
              const double wv = (double)(L.wz[k] * L.wy[j]) * s0;
#pragma unroll
              for (int i = 0; i < DP; ++i) atomicAdd(row + i, (C)fix1(L.wx[i], wv));
            }
          }
      } else {
        Axis<DP> ax, ay, az;
        constexpr int NEED = GRAD ? (kW | kDW) : kW;
        make_axis<DP, NEED>(x, DP, a.d0x, 1, chan_code(a, c, 0, a.bx), ax);
        make_axis<DP, NEED>(y, DP, a.d0y, a.d0x, chan_code(a, c, 1, a.by), ay);
        make_axis<DP, NEED>(z, DP, a.d0z, a.d0x * a.d0y, chan_code(a, c, 2, a.bz), az);
        if (GRAD) scatter_grad<DP>(dst, ax, ay, az, g0, g1, g2);
        else scatter_val<DP>(dst, ax, ay, az, g0);
      }
    }
    if (lds) {
      __syncthreads();
      flush_acc(cbox, S, scale, dst, a, chan_signs(a, c));
      __syncthreads();
    }
  }
  __syncthreads();  // the tile state may be reused by the next tile
}

// Waves per SIMD each 16x8x8 tile kernel is compiled for (one wave per SIMD per
// workgroup): as many workgroups as its LDS box lets share a CU, where the register
// budget that implies spills at most a few registers and measured faster; 0 leaves the
// compiler's choice.
template <int DP, int KIND>  // KIND: gather kinds 0-6, push 7, push_grad 8
constexpr int tile_waves() {
  return DP == 2 ? (KIND == 3 ? 5 : KIND == 7 ? 6 : 0)
       : DP == 3 ? (KIND == 0 || KIND == 4 ? 6 : KIND == 1 || KIND == 5 || KIND == 7 ? 5 : KIND == 2 || KIND == 3 ? 4 : 0)
       : DP == 4 ? (KIND == 2 ? 3 : KIND == 6 ? 4 : KIND == 7 || KIND == 0 ? 5 : 0)
                 : 0;
}

template <int DP, int TX, int TY, int TZ, int CAP, bool GRAD>
__global__ __launch_bounds__(kTB, (tile_waves<DP, GRAD ? 8 : 7>())) void k_push_tile(KArgs a) {
  __shared__ acc_t box[narrow_cells<GRAD>() ? (CAP + 1) / 2 : CAP];
  __shared__ TileState S;
  __shared__ float wsum[kTB / 64];
  push_tile<DP, TX, TY, TZ, CAP, GRAD>(a, xcd_block(blockIdx.x, gridDim.x), blockIdx.y, box, wsum, S);
}

// ---- gathers: pull (KIND 0), phi pullback of pull (1), fused objective (2), the ------
// ---- vector-Jacobian product of pull (3: push(delta) and the phi pullback together), -
// ---- pull_grad (4), pull_hess (5) and the pullback of pull_grad (6) ------------------
// The contractions run one axis at a time -- x within each row, then y within each
// plane, then z -- which takes fewer FMAs than weighting every row by its yz product
// (a cubic gradient: 188 instead of 224 per sample).
template <int DP>
__device__ __forceinline__ float lds_val(const float* box, const LSample<DP>& L, int BX, int BXY) {
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    float p = 0.f;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      const float* row = box + L.base + BX * j + BXY * k;
      float r = 0.f;
#pragma unroll
      for (int i = 0; i < DP; ++i) r += L.wx[i] * row[i];
      p += L.wy[j] * r;
    }
    acc += L.wz[k] * p;
  }
  return acc;
}

template <int DP, bool VAL>
__device__ __forceinline__ void lds_grad(const float* box, const LSample<DP>& L, int BX, int BXY, float& val,
                                         float& gx, float& gy, float& gz) {
  val = gx = gy = gz = 0.f;
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    float pv = 0.f, px = 0.f, py = 0.f;  // plane k: sum_j wy r, sum_j wy rd, sum_j dy r
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      const float* row = box + L.base + BX * j + BXY * k;
      float r = 0.f, rd = 0.f;
#pragma unroll
      for (int i = 0; i < DP; ++i) {
        const float v = row[i];
        r += L.wx[i] * v;
        rd += L.dx[i] * v;
      }
      pv += L.wy[j] * r;
      px += L.wy[j] * rd;
      py += L.dy[j] * r;
    }
    if (VAL) val += L.wz[k] * pv;
    gx += L.wz[k] * px;
    gy += L.wz[k] * py;
    gz += L.dz[k] * pv;
  }
}

// Hessian (xx, yy, zz, xy, xz, yz) from the box, one axis at a time as lds_grad.
template <int DP>
__device__ __forceinline__ void lds_hess(const float* box, const LSampleH<DP>& L, int BX, int BXY, float (&h)[6]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) h[q] = 0.f;
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    float pw = 0.f, pwd = 0.f, pwh = 0.f, pd = 0.f, pdd = 0.f, ph = 0.f;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      const float* row = box + L.base + BX * j + BXY * k;
      float r = 0.f, rd = 0.f, rh = 0.f;
#pragma unroll
      for (int i = 0; i < DP; ++i) {
        const float v = row[i];
        r += L.wx[i] * v;
        rd += L.dx[i] * v;
        rh += L.hx[i] * v;
      }
      pw += L.wy[j] * r;
      pwd += L.wy[j] * rd;
      pwh += L.wy[j] * rh;
      pd += L.dy[j] * r;
      pdd += L.dy[j] * rd;
      ph += L.hy[j] * r;
    }
    h[0] += L.wz[k] * pwh;
    h[1] += L.wz[k] * ph;
    h[2] += L.hz[k] * pw;
    h[3] += L.wz[k] * pdd;
    h[4] += L.dz[k] * pwd;
    h[5] += L.dz[k] * pd;
  }
}

// Gradient gather of one sample outside the tile's box, through global memory; for the
// objective (KIND 2) it also scatters the residual, for the VJP (3) the upstream gradient.
// Returns the weight w of the pullback (the upstream gradient, or the residual). Rarely
// taken, so it gathers lean: it must not set the register budget of the whole kernel.
template <int DP, int KIND>
__device__ __forceinline__ float outlier_grad(const KArgs& a, const float* __restrict__ src, float* __restrict__ dst2,
                                           float x, float y, float z, float aux, float& gx, float& gy, float& gz) {
  Axis<DP> ax, ay, az;
  make_axis<DP, kW | kDW>(x, DP, a.d0x, 1, a.bx, ax);
  make_axis<DP, kW | kDW>(y, DP, a.d0y, a.d0x, a.by, ay);
  make_axis<DP, kW | kDW>(z, DP, a.d0z, a.d0x * a.d0y, a.bz, az);
  float val = 0.f;
  if (KIND == 2) sample_val_grad<DP, true>(src, ax, ay, az, val, gx, gy, gz);
  else sample_grad<DP, true>(src, ax, ay, az, gx, gy, gz);
  const float w = (KIND == 2) ? val - aux : aux;
  if ((KIND == 2 || KIND == 3) && isfinite(w)) scatter_val<DP>(dst2, ax, ay, az, w);
  return w;
}

// LDS box of a gather kind, in floats. The scattering kinds' box first holds the staged
// image (float) and then, reused, the scatter accumulator (twice as large in 64-bit cells).
template <int CAP, int KIND>
constexpr int box_floats() {
  return (KIND == 2 || KIND == 3) && !narrow_cells<false>() ? 2 * CAP : CAP;
}

// One tile of pull (KIND 0), the pull pullback (1), the fused objective (2), the VJP
// of pull (3: dphi = sum_c delta_c grad f_c and df0 += push(delta) from one staging),
// pull_grad (4: the three components of grad f_c per channel), pull_hess (5: the 3x3
// Hessian per channel) or the pullback of pull_grad (6: dphi_b = sum_{a,c} delta_{a,c} H_{a,b,c}).
// box: box_floats<CAP, KIND>() floats (8-byte aligned); wsum: kTB/64 floats.
// The pullback of every sample is summed over channels in registers (gsum), and the
// objective keeps each sample's residual in registers between its two passes.
template <int DP, int TX, int TY, int TZ, int CAP, int KIND>
__device__ __forceinline__ void gather_tile(const KArgs& a, int tile, int bb, float* box, float* wsum,
                                            TileState& S) {
  constexpr int VPT = Shape<TX, TY, TZ>::VPT;
  constexpr bool AUX = KIND >= 1 && KIND <= 3;  // reads an upstream gradient or target per sample
  constexpr bool RED = AUX || KIND == 6;         // writes dphi, reduced over the channels
  const float* phi = a.phi + bb * a.phi_bs;
  float auxv[VPT];  // upstream gradient / target; channel 0's is loaded before phi is read
  auto prefetch = [&](const Vox<VPT>& W) {
#pragma unroll
    for (int q = 0; q < VPT; ++q)
      auxv[q] = (AUX && ((W.inside >> q) & 1u)) ? a.aux[bb * a.aux_bs + W.v[q]] : 0.f;
  };
  Vox<VPT> V;
  tile_setup<DP, TX, TY, TZ, CAP>(a, phi, tile, V, S, prefetch);
  const bool lds = S.lds != 0;
  const int BX = S.b[0], BXY = S.b[0] * S.b[1];
  const long long n1 = a.n1;
  float gsum[3][VPT];
#pragma unroll
  for (int q = 0; q < VPT; ++q) gsum[0][q] = gsum[1][q] = gsum[2][q] = 0.f;
  float lsum = 0.f;
  for (int c = 0; c < a.nchan; ++c) {
    const float* src = a.f + bb * a.f_bs + c * a.f_cs;
    // KIND 2/3: in-window samples with a finite weight, scattered via LDS; their weight
    // (the residual, which replaces the target in auxv, or the upstream gradient) is auxv
    unsigned rfin = 0u;
#pragma unroll
    for (int q = 0; q < VPT; ++q)
      if (c > 0)  // loaded before staging so both overlap
        auxv[q] = (AUX && ((V.valid >> q) & 1u)) ? a.aux[bb * a.aux_bs + c * a.aux_cs + V.v[q]] : 0.f;
    if (lds) {
      stage_box(box, S, src, a, chan_signs(a, c));
      __syncthreads();
    }
#pragma unroll 1
    for (int q = 0; q < VPT; ++q, V.rotate(), rot(auxv), rot_rows(gsum)) {
      if (!((V.inside >> q) & 1u)) continue;
      const int v = V.v[0];
      if (!((V.valid >> q) & 1u)) {
        if (KIND == 0) a.out[bb * a.out_bs + c * a.out_cs + v] = __builtin_nanf("");
        if (KIND == 4) put3(a.out + bb * a.out_bs + c * a.out_cs + v, n1, __builtin_nanf(""), __builtin_nanf(""),
                            __builtin_nanf(""), 0);
        if (KIND == 5)
          for (int e = 0; e < 9; ++e) a.out[bb * a.out_bs + c * a.out_cs + v + e * n1] = __builtin_nanf("");
        continue;
      }
      const float x = V.x[0], y = V.y[0], z = V.z[0];
      const bool use_lds = lds && ((V.inl >> q) & 1u);
      if (KIND == 0) {
        float val;
        if (use_lds) {
          LSample<DP> L;
          lsample<DP, false>(x, y, z, S, L);
          val = lds_val<DP>(box, L, BX, BXY);
        } else {
          Axis<DP> ax, ay, az;
          make_axis<DP, kW>(x, DP, a.d0x, 1, chan_code(a, c, 0, a.bx), ax);
          make_axis<DP, kW>(y, DP, a.d0y, a.d0x, chan_code(a, c, 1, a.by), ay);
          make_axis<DP, kW>(z, DP, a.d0z, a.d0x * a.d0y, chan_code(a, c, 2, a.bz), az);
          val = sample_val<DP, true>(src, ax, ay, az);
        }
        if (a.vscale != 1.0f || a.add_id) {
          const int l = threadIdx.x + kTB * q;
          val = pull_out(a, val, c, S.org[0] + l % TX, S.org[1] + (l / TX) % TY, S.org[2] + l / (TX * TY));
        }
        a.out[bb * a.out_bs + c * a.out_cs + v] = val;
        continue;
      }
      if (KIND == 5 || KIND == 6) {
        float h[6];
        if (use_lds) {
          LSampleH<DP> L;
          lsample_h<DP>(x, y, z, S, L);
          lds_hess<DP>(box, L, BX, BXY, h);
        } else {
          Axis<DP> ax, ay, az;
          make_axis<DP, kW | kDW | kHW>(x, DP, a.d0x, 1, a.bx, ax);
          make_axis<DP, kW | kDW | kHW>(y, DP, a.d0y, a.d0x, a.by, ay);
          make_axis<DP, kW | kDW | kHW>(z, DP, a.d0z, a.d0x * a.d0y, a.bz, az);
          sample_hess<DP>(src, ax, ay, az, h);
        }
        if (KIND == 5) {  // column-major 3x3 (pushpull.c:129-134)
          float* o = a.out + bb * a.out_bs + c * a.out_cs + v;
          o[0] = h[0]; o[4 * n1] = h[1]; o[8 * n1] = h[2];
          o[1 * n1] = h[3]; o[3 * n1] = h[3];
          o[2 * n1] = h[4]; o[6 * n1] = h[4];
          o[5 * n1] = h[5]; o[7 * n1] = h[5];
        } else {
          const float* dl = a.aux + bb * a.aux_bs + c * a.aux_cs + v;
          const float d0 = dl[0], d1 = dl[n1], d2 = dl[2 * n1];
          gsum[0][0] += d0 * h[0] + d1 * h[3] + d2 * h[4];
          gsum[1][0] += d0 * h[3] + d1 * h[1] + d2 * h[5];
          gsum[2][0] += d0 * h[4] + d1 * h[5] + d2 * h[2];
        }
        continue;
      }
      if (KIND == 4) {
        float gx, gy, gz;
        if (use_lds) {
          LSample<DP> L;
          lsample<DP, true>(x, y, z, S, L);
          float val;
          lds_grad<DP, false>(box, L, BX, BXY, val, gx, gy, gz);
        } else {
          Axis<DP> ax, ay, az;
          make_axis<DP, kW | kDW>(x, DP, a.d0x, 1, a.bx, ax);
          make_axis<DP, kW | kDW>(y, DP, a.d0y, a.d0x, a.by, ay);
          make_axis<DP, kW | kDW>(z, DP, a.d0z, a.d0x * a.d0y, a.bz, az);
          sample_grad<DP, true>(src, ax, ay, az, gx, gy, gz);
        }
        put3(a.out + bb * a.out_bs + c * a.out_cs + v, n1, gx, gy, gz, 0);
        continue;
      }
      const float aux = auxv[0];
      float val = 0.f, gx, gy, gz, w;
      if (use_lds) {
        LSample<DP> L;
        lsample<DP, true>(x, y, z, S, L);
        lds_grad<DP, KIND == 2>(box, L, BX, BXY, val, gx, gy, gz);
        w = (KIND == 2) ? val - aux : aux;
        if ((KIND == 2 || KIND == 3) && isfinite(w)) {
          if (KIND == 2) auxv[0] = w;
          rfin |= 1u << q;
        }
      } else {
        w = outlier_grad<DP, KIND>(a, src, a.out2 + bb * a.out2_bs + c * a.out2_cs, x, y, z, aux, gx, gy, gz);
      }
      if (KIND == 2 && isfinite(w)) lsum += 0.5f * w * w;
      gsum[0][0] += w * gx;
      gsum[1][0] += w * gy;
      gsum[2][0] += w * gz;
    }
    if (lds) {
      __syncthreads();
      if (KIND == 2 || KIND == 3) {  // scatter the residuals (gradients) of the in-window samples via the fixed-point box
        constexpr bool NW = narrow_cells<false>();
        typedef typename CellOf<NW>::T C;
        float vb = 0.f;
#pragma unroll
        for (int q = 0; q < VPT; ++q)
          if ((rfin >> q) & 1u) {
            const float r = fabsf(auxv[q]);
            vb = NW ? vb + r : fmaxf(vb, r);
          }
        const float m = block_max<NW>(vb, wsum);  // also orders the staging reads
        const double scale = NW ? fixed_scale32<DP>(m) : fixed_scale(m);
        C* cbox = reinterpret_cast<C*>(box);
        zero_acc(cbox, S.cells);
        __syncthreads();
#pragma unroll 1
        for (int q = 0; q < VPT; ++q, V.rotate(), rot(auxv)) {
          if (!((rfin >> q) & 1u)) continue;
          LSample<DP> L;
          lsample<DP, false>(V.x[0], V.y[0], V.z[0], S, L);
          const double ws = (double)auxv[0] * scale;
#pragma unroll
          for (int k = 0; k < DP; ++k)
#pragma unroll
            for (int j = 0; j < DP; ++j) {
              C* row = cbox + L.base + BX * j + BXY * k;
              const double wv = (double)(L.wz[k] * L.wy[j]) * ws;
#pragma unroll
              for (int i = 0; i < DP; ++i) atomicAdd(row + i, (C)fix1(L.wx[i], wv));
            }
        }
        __syncthreads();
        flush_acc(cbox, S, scale, a.out2 + bb * a.out2_bs + c * a.out2_cs, a, chan_signs(a, c));
        __syncthreads();
      }
    }
  }
  if (RED) {
    const float nan = __builtin_nanf("");
#pragma unroll 1
    for (int q = 0; q < VPT; ++q, V.rotate(), rot_rows(gsum)) {
      if (!((V.inside >> q) & 1u)) continue;
      const int v = V.v[0];
      if ((V.valid >> q) & 1u) put3(a.out + bb * a.out_bs + v, n1, gsum[0][0], gsum[1][0], gsum[2][0], a.accum);
      else put3(a.out + bb * a.out_bs + v, n1, nan, nan, nan, 0);
    }
  }
  if (KIND == 2) {
    lsum = __ockl_wfred_add_f32(lsum);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kTB / 64; ++w) t += (double)wsum[w];
      unsafeAtomicAdd(a.red, t);
    }
  }
  __syncthreads();  // the tile state and boxes may be reused by the next tile
}

template <int DP, int TX, int TY, int TZ, int CAP, int KIND>
__global__ __launch_bounds__(kTB, (tile_waves<DP, KIND>())) void k_gather_tile(KArgs a) {
  __shared__ __align__(8) float box[box_floats<CAP, KIND>()];
  __shared__ TileState S;
  __shared__ float wsum[kTB / 64];
  gather_tile<DP, TX, TY, TZ, CAP, KIND>(a, xcd_block(blockIdx.x, gridDim.x), blockIdx.y, box, wsum, S);
}

// ---- launcher ------------------------------------------------------------------
// Box budgets in cells, per degree: small boxes let more workgroups share a CU, whose
// latency chains (load phi, size the box, stage, compute) then overlap. A tile whose box
// exceeds the budget runs per sample through global memory. Gathers stage floats;
// scatters and the objective accumulate fixed-point cells (32-bit, push_grad 64-bit).
template <int DP, int KIND>
constexpr int cap_cells() {
  // (the cubic pull: 7680 floats, so that five workgroups of 91 registers share a CU)
  return KIND == 0 && DP == 4 ? 7680 : (KIND == 0 || KIND == 1) ? (DP <= 2 ? 4096 : DP == 3 ? 6400 : 9472)
                                  : (DP <= 2 ? 3200 : DP == 3 ? 4800 : 6400);  // DP 5 as DP 4
}

template <int TX, int TY, int TZ>
static int tiles_of(const KArgs& a) {
  return ((a.d1x + TX - 1) / TX) * ((a.d1y + TY - 1) / TY) * ((a.d1z + TZ - 1) / TZ);
}

template <int DP, int TX, int TY, int TZ>
static hipError_t launch_shape(Op op, const KArgs& a, int nbatch, hipStream_t s) {
  const dim3 grid((unsigned)tiles_of<TX, TY, TZ>(a), (unsigned)nbatch, 1), block(kTB);
  switch (op) {
    case OP_PULL:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 0>(), 0>), grid, block, 0, s, a);
      break;
    case OP_PULLG:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 0>(), 4>), grid, block, 0, s, a);
      break;
    case OP_PULLH:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 0>(), 5>), grid, block, 0, s, a);
      break;
    case OP_PULLG_BWD:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 1>(), 6>), grid, block, 0, s, a);
      break;
    case OP_PULL_BWD:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 1>(), 1>), grid, block, 0, s, a);
      break;
    case OP_OBJGRAD:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 2>(), 2>), grid, block, 0, s, a);
      break;
    case OP_PULL_VJP:
      hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, cap_cells<DP, 2>(), 3>), grid, block, 0, s, a);
      break;
    case OP_PUSH:
      hipLaunchKernelGGL((k_push_tile<DP, TX, TY, TZ, cap_cells<DP, 3>(), false>), grid, block, 0, s, a);
      break;
    case OP_PUSHG:
      hipLaunchKernelGGL((k_push_tile<DP, TX, TY, TZ, cap_cells<DP, 4>(), true>), grid, block, 0, s, a);
      break;
    default: return hipErrorNotSupported;
  }
  return hipGetLastError();
}

template <int DP>
static hipError_t launch_deg(Op op, const KArgs& a, int nbatch, hipStream_t s, TileCfg cfg) {
  switch (cfg.shape) {
    case 1: return launch_shape<DP, 32, 8, 4>(op, a, nbatch, s);
    case 2: return launch_shape<DP, 16, 16, 4>(op, a, nbatch, s);
    default: return launch_shape<DP, 16, 8, 8>(op, a, nbatch, s);
  }
}

// ---- batched launch over volumes of uneven size ----------------------------------------
// A plan orders its volumes by degree; each degree present gets one launch of a kernel
// compiled for that degree only (a single kernel switching over five degrees would carry
// the register footprint of all of them). Within a launch the tiles of the group's volumes
// are numbered consecutively (BatchVol::tile0, relative to the group) and cut into eight
// contiguous ranges of equal work, one per XCD (workgroup b runs on XCD b % 8), so that
// neighbouring tiles share an L2 and no XCD idles while another still holds the heavy
// (many-channel) volumes. Degrees 1-4 use the LDS tile bodies; degree 0 runs the same
// bodies with an empty box budget, i.e. per sample through global memory.
constexpr int kBTX = 16, kBTY = 8, kBTZ = 8;

// k_batch kinds: gather kinds 0-2, push 3, push_grad 4, the pull VJP 5 (gather kind 3)
constexpr int batch_gather_kind(int kind) { return kind <= 2 ? kind : kind == 5 ? 3 : -1; }
constexpr int batch_waves_kind(int kind) { return kind <= 2 ? kind : kind == 5 ? 3 : kind + 4; }

template <int DP, int KIND>
__global__ __launch_bounds__(kTB, (tile_waves<DP, batch_waves_kind(KIND)>())) void k_batch(
    const BatchVol* __restrict__ vols, int nvol, XcdSplit split) {
  constexpr int KG = batch_gather_kind(KIND) >= 0 ? batch_gather_kind(KIND) : 0;
  constexpr bool GATHER = batch_gather_kind(KIND) >= 0;
  constexpr int CAPE = (DP >= 2 && DP <= 5) ? cap_cells<DP, KIND>() : 0;
  constexpr int BOXF = GATHER ? box_floats<CAPE, KG>() : narrow_cells<KIND == 4>() ? CAPE : 2 * CAPE;  // push cells
  __shared__ __align__(8) float box[BOXF > 0 ? BOXF : 2];
  __shared__ TileState S;
  __shared__ float wsum[kTB / 64];
  // workgroup b runs on XCD b % 8 and takes the next tile of that XCD's range
  const int x = blockIdx.x % kBatchXcds;
  const int t = split.s[x] + (int)(blockIdx.x / kBatchXcds);
  if (t >= split.s[x + 1]) return;
  int lo = 0, hi = nvol - 1;  // last volume whose first tile is <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vols[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  const KArgs a = vols[lo].a;
  const int tile = t - vols[lo].tile0;
  if (GATHER) gather_tile<DP, kBTX, kBTY, kBTZ, CAPE, KG>(a, tile, 0, box, wsum, S);
  else push_tile<DP, kBTX, kBTY, kBTZ, CAPE, KIND == 4>(a, tile, 0, reinterpret_cast<acc_t*>(box), wsum, S);
}

int batch_tiles(const KArgs& a) { return tiles_of<kBTX, kBTY, kBTZ>(a); }

// ---- LDS atomic calibration --------------------------------------------------------
// The rate of LDS integer adds on cells of type T (the cubic value scatter's: 32-bit)
// under the scatter's own access pattern, the on-chip bound of push and the VJP: every lane adds the 4x4x4 taps of cubic samples into a box shaped like
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
  hipLaunchKernelGGL((k_lds_add_rate<typename CellOf<narrow_cells<false>()>::T>), dim3((unsigned)blocks), dim3(kTB),
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

template <int KIND>
static const void* batch_kernel(int dp) {
  switch (dp) {
    case 1: return reinterpret_cast<const void*>(&k_batch<1, KIND>);
    case 2: return reinterpret_cast<const void*>(&k_batch<2, KIND>);
    case 3: return reinterpret_cast<const void*>(&k_batch<3, KIND>);
    case 4: return reinterpret_cast<const void*>(&k_batch<4, KIND>);
    default: return reinterpret_cast<const void*>(&k_batch<5, KIND>);
  }
}

static const void* batch_kernel_for(Op op, int dp) {
  switch (op) {
    case OP_PULL: return batch_kernel<0>(dp);
    case OP_PULL_BWD: return batch_kernel<1>(dp);
    case OP_OBJGRAD: return batch_kernel<2>(dp);
    case OP_PULL_VJP: return batch_kernel<5>(dp);
    case OP_PUSH: return batch_kernel<3>(dp);
    case OP_PUSHG: return batch_kernel<4>(dp);
    default: return nullptr;
  }
}

hipError_t launch_batch_zero(Op op, const BatchVol* dvols, int nvol, hipStream_t s) {
  if (nvol <= 0 || !(op == OP_PUSH || op == OP_PUSHG || op == OP_OBJGRAD || op == OP_PULL_VJP)) return hipSuccess;
  hipLaunchKernelGGL(k_batch_zero, dim3(512, (unsigned)nvol), dim3(256), 0, s, dvols, (int)op);
  return hipGetLastError();
}

hipError_t launch_batch_group(Op op, int dp, const BatchVol* dvols, int nvol, const XcdSplit& split, hipStream_t s) {
  int longest = 0;
  for (int x = 0; x < kBatchXcds; ++x) longest = std::max(longest, split.s[x + 1] - split.s[x]);
  if (nvol <= 0 || longest <= 0) return hipSuccess;
  const void* k = batch_kernel_for(op, dp);
  if (!k) return hipErrorNotSupported;
  XcdSplit sp = split;
  void* args[] = {(void*)&dvols, (void*)&nvol, (void*)&sp};
  const long long grid = (long long)longest * kBatchXcds;
  if (grid > INT_MAX) return hipErrorInvalidValue;
  return hipLaunchKernel(k, dim3((unsigned)grid), dim3(kTB), args, 0, s);
}

hipError_t launch_tiled(Op op, const KArgs& a, int nbatch, hipStream_t s, TileCfg cfg) {
  if (op != OP_PULL && op != OP_PULLG && op != OP_PULLH && op != OP_PULL_BWD && op != OP_PULLG_BWD &&
      op != OP_OBJGRAD && op != OP_PULL_VJP && op != OP_PUSH && op != OP_PUSHG)
    return hipErrorNotSupported;
  if (a.n1 <= 0 || nbatch <= 0) return hipSuccess;
  if (a.dpx != a.dpy || a.dpx != a.dpz) return hipErrorNotSupported;
  if ((long long)a.d1x * a.d1y * a.d1z != a.n1) return hipErrorNotSupported;
  switch (a.dpx) {
    case 2: return launch_deg<2>(op, a, nbatch, s, cfg);
    case 3: return launch_deg<3>(op, a, nbatch, s, cfg);
    case 4: return launch_deg<4>(op, a, nbatch, s, cfg);
    case 5: return launch_deg<5>(op, a, nbatch, s, cfg);
    default: return hipErrorNotSupported;
  }
}

}  // namespace pp
