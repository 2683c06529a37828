// pp_tiled_impl.hpp -- the tiled kernels' device code and per-shape launchers, shared by
// pp_tiled.hip (isotropic degrees) and pp_tiled_aniso.hip (a degree per axis), each of which
// instantiates its own single-call and batch kernels.
#pragma once
#define PP_KERNEL_SOURCE 1  // kernel arguments carry the global address space (pp_kernels.hpp)
#include "pp_sample.hpp"

#include <algorithm>
#include <climits>
#include <type_traits>

namespace pp {

constexpr int kTB = 256;

// A tile kernel's LDS box is the launch's dynamic LDS: `cap` floats (8-byte aligned) that
// hold the staged image or the fixed-point accumulator, and at their end the row table.
extern __shared__ __align__(16) float pp_box[];

// Largest box extent along y or z: 2 * (window half-width) + dp for every tile shape.
constexpr int kMaxYZ = 64;

struct TileState {
  int lo[3];      // unfolded image index of box cell (0,0,0), per axis
  int b[3];       // box extent per axis
  int Pf, Qf;     // LDS pitches of a box row and plane: staged image (floats)
  int Pc, Qc;     // ... and fixed-point accumulator (64-bit cells); see tile_setup
  int cells_f;    // floats of the staged image: Qf * b[2]
  int cells_c;    // 64-bit cells of the accumulator: Qc * b[2]
  int lds;        // 1 when the box is used
  int vec;        // 1 when the staged rows move in 16-byte chunks (see tile_setup)
  int nonfin;     // anisotropic degrees: 1 + the last channel whose staged box held a non-finite value
  float ctr[3];   // window centre: the tile's centre sample (else the tile centre, mapped)
  int org[3];     // tile origin in the phi grid
  int red[6][4];  // per-wave partial min/max of the in-window knots
  float redf[4];  // per-wave partial bound (tile_setup's `bound` hook)
  int rows;       // box rows: b[1] * b[2]
  int tab;        // float offset, in the launch's box memory, of the row table (RowEnt[rows])
};

// One box row (plane indices ry, rz; row r = ry + b[1] rz): the LDS index of its cell 0
// (ry P + rz Q, in the layout's units) with the boundary's sign flips of the row's y and
// z planes in bits 31 and 30, and the folded image offset of its column 0 before the x
// fold (d0x fold(lo[1] + ry) + d0x d0y fold(lo[2] + rz)). Built once per tile, it makes
// staging and flushing a cell a table read and a few instructions.
struct RowEnt {
  int cell;
  int goff;
};
constexpr int kFlipY = INT_MIN, kFlipZ = 1 << 30, kCellMask = (1 << 30) - 1;

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
  // local index of the centre sample (tile_pos order)
  static constexpr int LC = TX / 2 + TX * (TZ / 2 + TZ * (TY / 2));
};

// Position in the tile of local sample l = tid + kTB * q (q: the thread's sample slot):
// x fastest, then z, then y. A wave's 64 lanes then hold TX consecutive x positions on
// 64 / TX consecutive z planes of one y row, and a thread's slots step along y. With a
// box plane pitch Q = TX (mod 2 TX) the lanes of one LDS read fall on distinct banks (the
// 64 banks of gfx950) for samples whose knots follow their positions.
template <int TX, int TY, int TZ>
__device__ __forceinline__ void tile_pos(int l, int& x, int& y, int& z) {
  x = l % TX;
  z = (l / TX) % TZ;
  y = l / (TX * TZ);
}

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

// A workgroup-uniform value read from LDS, moved to a scalar register: what is computed
// from it (the box sizing) then runs on the scalar unit instead of in every lane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// floor(a / b) for 0 <= a < 2^16 and 1 <= b <= 256 through a float reciprocal: (a + 0.5) / b
// lies at least 0.5 / b from an integer, far beyond the few ulps of error of the product,
// so the conversion truncates to the exact quotient (the generic integer division takes
// some thirty instructions).
__device__ __forceinline__ int small_div(int a, int b) {
  return (int)(((float)a + 0.5f) * __builtin_amdgcn_rcpf((float)b));
}

// Wave-wide reductions (DPP lane moves, not LDS-routed permutes).
// (integer min/max fuse each DPP step into one instruction; the float forms take four)
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }

// A window centre at most this far from the origin keeps every in-window coordinate
// within kFastLimit (the range of knot_fast and the box weights).
constexpr float kCentreLimit = 0.5f * kFastLimit;

template <int DP>
__device__ __forceinline__ int knot_fast(float x) {
  return (int)ceilf(x - 0.5f * (float)DP);
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
//   bound()   a per-thread value >= 0 whose workgroup maximum lands in *block_bound;
//   zero_box  when given, the box's 64-bit cells are zeroed before the final barrier.
// The box is used when the layouts the op needs (GF: the staged image, GC: the
// accumulator) fit the cap floats of LDS the launch provides.
template <int DP, int TX, int TY, int TZ, bool GF, bool GC, class Pre = NoPrefetch, class Bound = NoBound>
__device__ __forceinline__ void tile_setup(const KArgs& a, const float* __restrict__ phi, int tile, int cap,
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
  // A thread's slots share their x and z positions and step along y by DY rows (tile_pos of
  // tid + kTB q), so a slot's voxel index is the first slot's plus q DY d1x.
  static_assert(kTB % (TX * TZ) == 0 && (kTB / TX) % TZ == 0, "slots differ in y only");
  constexpr int DY = kTB / (TX * TZ);
  int gx, gy0, gz;
  tile_pos<TX, TY, TZ>(tid, gx, gy0, gz);
  gx += x0;
  gy0 += y0;
  gz += z0;
  const bool inxz = gx < a.d1x && gz < a.d1z;
  const int v0 = gx + a.d1x * (gy0 + a.d1y * gz), dv = DY * a.d1x;
#pragma unroll
  for (int q = 0; q < VPT; ++q) {
    const bool in = inxz && gy0 + DY * q < a.d1y;
    V.v[q] = in ? v0 + q * dv : 0;
    if (in) V.inside |= 1u << q;
  }
  pre(V);
  // Every slot's deformation values are loaded before any is used (a slot outside the
  // grid reads voxel 0, a valid address, and is masked below): loads under a per-slot
  // condition would each be waited for before the next slot's were issued, one memory
  // round trip per slot.
  float px[VPT], py[VPT], pz[VPT];
  if (!a.affine) {
    const long long n1 = a.n1;
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      px[q] = phi[V.v[q]];
      py[q] = phi[V.v[q] + n1];
      pz[q] = phi[V.v[q] + 2 * n1];
    }
  }
#pragma unroll
  for (int q = 0; q < VPT; ++q) {
    const int gy = gy0 + DY * q;
    float x = 0.f, y = 0.f, z = 0.f;
    bool ok = false;
    if ((V.inside >> q) & 1u) {
      if (a.affine) affine_x(a.aff, gx, gy, gz, x, y, z);
      else coords_from(a, px[q], py[q], pz[q], gx, gy, gz, x, y, z);
      ok = voxok(x, y, z, a.ext, a.lx, a.ly, a.lz);
    }
    V.x[q] = x;
    V.y[q] = y;
    V.z[q] = z;
    if (ok) V.valid |= 1u << q;
    // window centre: the centre sample, else the tile centre mapped by the grid-size ratio
    if (q == SH::LC / kTB && tid == SH::LC % kTB) {
      S.org[0] = x0;
      S.org[1] = y0;
      S.org[2] = z0;
      if (ok && fabsf(x) < kCentreLimit && fabsf(y) < kCentreLimit && fabsf(z) < kCentreLimit) {
        S.ctr[0] = x;
        S.ctr[1] = y;
        S.ctr[2] = z;
      } else {  // (double: no 64-bit division)
        S.ctr[0] = fminf((float)((double)(x0 + TX / 2) * a.d0x / a.d1x), kCentreLimit);
        S.ctr[1] = fminf((float)((double)(y0 + TY / 2) * a.d0y / a.d1y), kCentreLimit);
        S.ctr[2] = fminf((float)((double)(z0 + TZ / 2) * a.d0z / a.d1z), kCentreLimit);
      }
    }
  }
  __syncthreads();
  // In-window samples: valid, and within the half-widths of the centre on every axis (a
  // NaN coordinate fails the test). The box spans the knots of their extreme coordinates
  // (the knot is monotonic in the coordinate): each lane takes the extremes of its slots'
  // coordinates and only their knots are computed, then reduced as integers.
  const float cx = uni(S.ctr[0]), cy = uni(S.ctr[1]), cz = uni(S.ctr[2]);
  float mn0 = INFINITY, mn1 = INFINITY, mn2 = INFINITY, mx0 = -INFINITY, mx1 = -INFINITY, mx2 = -INFINITY;
#pragma unroll
  for (int q = 0; q < VPT; ++q) {
    const float x = V.x[q], y = V.y[q], z = V.z[q];
    if (!((V.valid >> q) & 1u) || !(fabsf(x - cx) <= (float)SH::WX) || !(fabsf(y - cy) <= (float)SH::WY) ||
        !(fabsf(z - cz) <= (float)SH::WZ))
      continue;
    V.inl |= 1u << q;
    mn0 = fminf(mn0, x); mx0 = fmaxf(mx0, x);
    mn1 = fminf(mn1, y); mx1 = fmaxf(mx1, y);
    mn2 = fminf(mn2, z); mx2 = fmaxf(mx2, z);
  }
  const bool lany = mn0 <= mx0;  // (a lane without an in-window sample contributes nothing)
  int kn0 = lany ? knot_fast<DP>(mn0) : INT_MAX, kn1 = lany ? knot_fast<DP>(mn1) : INT_MAX;
  int kn2 = lany ? knot_fast<DP>(mn2) : INT_MAX, kx0 = lany ? knot_fast<DP>(mx0) : INT_MIN;
  int kx1 = lany ? knot_fast<DP>(mx1) : INT_MIN, kx2 = lany ? knot_fast<DP>(mx2) : INT_MIN;
  kn0 = wave_min(kn0); kn1 = wave_min(kn1); kn2 = wave_min(kn2);
  kx0 = wave_max(kx0); kx1 = wave_max(kx1); kx2 = wave_max(kx2);
  float bnd = bound();
  if (block_bound)
    bnd = __ockl_wfred_max_f32(bnd);
  const int wid = tid >> 6;
  if ((tid & 63) == 0) {
    S.red[0][wid] = kn0; S.red[1][wid] = kn1; S.red[2][wid] = kn2;
    S.red[3][wid] = kx0; S.red[4][wid] = kx1; S.red[5][wid] = kx2;
    S.redf[wid] = bnd;
  }
  __syncthreads();
  if (block_bound) {
    float m = S.redf[0];
#pragma unroll
    for (int w = 1; w < kTB / 64; ++w) m = fmaxf(m, S.redf[w]);
    *block_bound = m;
  }
  // every thread sizes the box (a broadcast read of the partials)
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
  for (int w = 0; w < kTB / 64; ++w)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      lo[q] = min(lo[q], uni(S.red[q][w]));
      hi[q] = max(hi[q], uni(S.red[3 + q][w]));
    }
  bool any = lo[0] <= hi[0];  // some in-window sample
  if (!any) {
#pragma unroll
    for (int q = 0; q < 3; ++q) lo[q] = hi[q] = 0;
  }
  // LDS layouts: cell (i, j, k) at i + P j + Q k with P = bx. Plane pitches are padded to
  // Q = TX (mod 2 TX), so that the TX x-lanes on 64 / TX planes of a wave (tile_pos) fall
  // on distinct banks.
  auto pitch = [](int P, int by) { return P * by + ((TX - P * by) & (2 * TX - 1)); };
  auto fits = [&](int P, int by, int bz) {  // box cells (row pitch P) and the row table (two floats a row)
    const int q = pitch(P, by) * bz, t = 2 * by * bz;
    return (!GF || q + t <= cap) && (!GC || 2 * q + t <= cap);
  };
  // A box beyond the budget is clipped, one knot plane at a time on the side farther from
  // the window centre along its most spread axis, until it fits; the samples whose
  // footprint then leaves the box take the global path. (Only a box that cannot shrink
  // to the budget sends the whole tile there.)
  bool clipped = false;
  if (any && !fits(hi[0] - lo[0] + DP, hi[1] - lo[1] + DP, hi[2] - lo[2] + DP)) {
    const int c0[3] = {knot_fast<DP>(cx), knot_fast<DP>(cy), knot_fast<DP>(cz)};
    const int span[3] = {TX, TY, TZ};
    while (!fits(hi[0] - lo[0] + DP, hi[1] - lo[1] + DP, hi[2] - lo[2] + DP)) {
      int ax = -1;
      float worst = 0.f;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float r = (float)(hi[q] - lo[q] + 1) / (float)span[q];
        if (hi[q] > lo[q] && r > worst) {
          worst = r;
          ax = q;
        }
      }
      if (ax < 0) break;
      if (c0[ax] - lo[ax] > hi[ax] - c0[ax]) ++lo[ax];
      else --hi[ax];
      clipped = true;
    }
    any = fits(hi[0] - lo[0] + DP, hi[1] - lo[1] + DP, hi[2] - lo[2] + DP);
  }
  if (clipped) {  // (a uniform branch: every thread sized the same box)
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      if (!((V.inl >> q) & 1u)) continue;
      const int ox = knot_fast<DP>(V.x[q]), oy = knot_fast<DP>(V.y[q]), oz = knot_fast<DP>(V.z[q]);
      if (ox < lo[0] || ox > hi[0] || oy < lo[1] || oy > hi[1] || oz < lo[2] || oz > hi[2]) V.inl &= ~(1u << q);
    }
  }
  // Staging in 16-byte chunks: when the box lies inside the image along x (no x fold) and
  // image rows start on 16-byte boundaries (d0x, the channel and batch strides multiples of
  // four floats, an aligned base), the box is extended on its low x side to a multiple of
  // four (at most three columns, never read by a sample) and its row pitch padded to a
  // multiple of four, so that one 16-byte load and one 16-byte LDS store move four cells.
  bool vec = false;
  if (GF && any && lo[0] >= 0 && hi[0] + DP <= a.d0x && (a.d0x & 3) == 0 && ((a.f_cs | a.f_bs) & 3) == 0 &&
      (reinterpret_cast<unsigned long long>(a.f) & 15) == 0) {
    const int l0 = lo[0] & ~3;
    if (fits((hi[0] + DP - l0 + 3) & ~3, hi[1] - lo[1] + DP, hi[2] - lo[2] + DP)) {
      vec = true;
      lo[0] = l0;
    }
  }
  const int bx = hi[0] - lo[0] + DP, by = hi[1] - lo[1] + DP, bz = hi[2] - lo[2] + DP;
  const int Pf = vec ? (bx + 3) & ~3 : bx, Pc = Pf, Qf = pitch(Pf, by), Qc = Qf;
  const int cells_f = any ? Qf * bz : 0, cells_c = any ? Qc * bz : 0;
  const bool use = any;
  const int rows = by * bz, tab = (cap & ~1) - 2 * rows;
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      S.lo[q] = lo[q];
      S.b[q] = hi[q] - lo[q] + DP;
    }
    S.Pf = Pf;
    S.Qf = Qf;
    S.Pc = Pc;
    S.Qc = Qc;
    S.cells_f = cells_f;
    S.cells_c = cells_c;
    S.lds = use;
    S.vec = vec;
    S.nonfin = 0;
    S.rows = rows;
    S.tab = tab;
  }
  if (use && zero_box)
    for (int i = tid; i < cells_c; i += kTB) zero_box[i] = 0ull;
  if (use) {  // the row table, at the end of the box memory
    RowEnt* rt = reinterpret_cast<RowEnt*>(pp_box + tab);
    const float inv_by = 1.0f / (float)by;
    for (int r = tid; r < rows; r += kTB) {
      const int rz = (int)(((float)r + 0.5f) * inv_by), ry = r - rz * by;  // (exact: r < 2^14)
      bool ny, nz;
      const int fy = fold32(a.by, lo[1] + ry, a.d0y, ny), fz = fold32(a.bz, lo[2] + rz, a.d0z, nz);
      RowEnt e;
      e.cell = (ry * Pf + rz * Qf) | (ny ? kFlipY : 0) | (nz ? kFlipZ : 0);
      e.goff = a.d0x * (fy + a.d0y * fz);
      rt[r] = e;
    }
  }
  __syncthreads();
}

// The box is walked by the whole workgroup: thread t owns column c = t % b[0] of rows
// t / b[0], t / b[0] + RP, ... (RP = kTB / b[0] rows per pass; b[0] <= 2 WX + DP < kTB).
// A cell's LDS index and image offset come from its row's table entry and the column's
// x fold; kBatch cells' reads are issued before any is used.
constexpr int kBatch = 8;

struct RowWalk {
  int c, r, rp;  // column, first row, rows per pass
  int fx;        // folded x offset of the column
  int xflip;     // kFlipY when the column's x fold flips the sign (merged into the row parity)
  bool active;   // this thread owns a column
};

__device__ __forceinline__ RowWalk row_walk(const TileState& S, const KArgs& a) {
  const int BX = uni(S.b[0]);
  RowWalk w;
  w.r = small_div(threadIdx.x, BX);
  w.c = threadIdx.x - w.r * BX;
  w.rp = small_div(kTB, BX);
  w.active = w.r < w.rp;
  bool nx;
  w.fx = fold32(a.bx, S.lo[0] + w.c, a.d0x, nx);
  w.xflip = nx ? kFlipY : 0;
  return w;
}

// Sign of a cell: the parity of the flips of its row's y and z planes and of its column,
// each counted only for the axes (bits of `signs`) whose boundary sign applies.
__device__ __forceinline__ bool cell_neg(int cellword, const RowWalk& w, unsigned signs) {
  const bool fy = (signs & 2u) && (cellword & kFlipY);
  const bool fz = (signs & 4u) && (cellword & kFlipZ);
  const bool fx = (signs & 1u) && w.xflip;
  return fy != fz != fx;
}

// Stage the box from image channel `src` (fold + boundary sign applied here). FLIPS: some
// axis folds with code 2 (the only code whose mirrored copies change sign).
template <bool FLIPS>
__device__ __forceinline__ void stage_rows(float* box, const TileState& S, const float* __restrict__ src,
                                           unsigned signs, RowWalk w) {
  const RowEnt* rt = reinterpret_cast<const RowEnt*>(pp_box + S.tab);
  const int rows = S.rows;
  for (; w.r < rows; w.r += kBatch * w.rp) {
    RowEnt e[kBatch];
    float v[kBatch];
    // table entries are read unconditionally, rows past the end reading the last row's
    // (a read under a per-row condition is waited for before the next one is issued)
#pragma unroll
    for (int u = 0; u < kBatch; ++u) e[u] = rt[min(w.r + u * w.rp, rows - 1)];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = w.r + u * w.rp < rows ? src[e[u].goff + w.fx] : 0.f;
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (w.r + u * w.rp < rows)
        box[(e[u].cell & kCellMask) + w.c] = (FLIPS && cell_neg(e[u].cell, w, signs)) ? -v[u] : v[u];
  }
}

// Staging in 16-byte chunks (the box's vec layout: no x fold, rows and the box's x origin
// aligned to four floats): thread t moves chunk t % nc of rows t / nc, t / nc + rp, ...
// (nc chunks a row, rp = kTB / nc rows a pass); the chunks past the box's last column
// fill the row's padding with image cells that exist (the box ends inside the row).
constexpr int kVBatch = 4;  // rows a thread loads before storing (as many registers as kBatch floats)

template <bool FLIPS>
__device__ __forceinline__ void stage_rows_vec(float* box, const TileState& S, const float* __restrict__ src,
                                               const KArgs& a, unsigned signs) {
  const int nc = uni(S.Pf) >> 2, rows = uni(S.rows);
  const int rp = small_div(kTB, nc);
  int r = small_div(threadIdx.x, nc);
  const int c = threadIdx.x - r * nc;
  if (r >= rp) return;
  const RowEnt* rt = reinterpret_cast<const RowEnt*>(pp_box + uni(S.tab));
  const float* s0 = src + uni(S.lo[0]) + 4 * c;
  for (; r < rows; r += kVBatch * rp) {
    RowEnt e[kVBatch];
    float4 v[kVBatch];
#pragma unroll
    for (int u = 0; u < kVBatch; ++u) e[u] = rt[min(r + u * rp, rows - 1)];
#pragma unroll
    for (int u = 0; u < kVBatch; ++u)
      if (r + u * rp < rows) v[u] = *reinterpret_cast<const float4*>(s0 + e[u].goff);
#pragma unroll
    for (int u = 0; u < kVBatch; ++u)
      if (r + u * rp < rows) {
        float4 t = v[u];
        if (FLIPS && (((signs & 2u) && (e[u].cell & kFlipY)) != ((signs & 4u) && (e[u].cell & kFlipZ)))) {
          t.x = -t.x;
          t.y = -t.y;
          t.z = -t.z;
          t.w = -t.w;
        }
        *reinterpret_cast<float4*>(box + (e[u].cell & kCellMask) + 4 * c) = t;
      }
  }
}

// Anisotropic degrees: the box's taps outside an axis's support are weighted 0, and 0
// times a non-finite value is NaN where the reference, which never reads that tap, has a
// value. Non-finite staged values are therefore replaced by 0 here (thread t checks column
// t % P of rows t / P, t / P + kTB / P, ..., P the staged columns: the row pitch in the
// 16-byte layout, whose padding is staged too, else the box width), and a channel whose box
// held one takes the global path for all its samples, which reads each axis's own taps
// only. Returns whether this thread found one.
__device__ __forceinline__ bool sanitize_box(float* box, const TileState& S) {
  const int P = uni(S.vec) ? uni(S.Pf) : uni(S.b[0]), rows = uni(S.rows);
  const RowEnt* rt = reinterpret_cast<const RowEnt*>(pp_box + uni(S.tab));
  const int r0 = small_div(threadIdx.x, P), c = threadIdx.x - r0 * P, rp = small_div(kTB, P);
  bool bad = false;
  if (r0 < rp)
    for (int r = r0; r < rows; r += rp) {
      float* q = box + (rt[r].cell & kCellMask) + c;
      if (!isfinite(*q)) {
        *q = 0.f;
        bad = true;
      }
    }
  return bad;
}

__device__ __forceinline__ void stage_box(float* box, const TileState& S, const float* __restrict__ src,
                                          const KArgs& a, unsigned signs) {
  if (uni(S.vec)) {
    if (a.by == 2 || a.bz == 2) stage_rows_vec<true>(box, S, src, a, signs);
    else stage_rows_vec<false>(box, S, src, a, signs);
    return;
  }
  RowWalk w = row_walk(S, a);
  if (!w.active) return;
  if (a.bx == 2 || a.by == 2 || a.bz == 2) stage_rows<true>(box, S, src, signs, w);
  else stage_rows<false>(box, S, src, signs, w);
}

// ---- fixed-point scatter accumulation ---------------------------------------------
// gfx950 executes float atomics (LDS ds_add_f32 and global alike) at a small fraction
// of the rate of integer LDS atomics (ds_add_f32: ~190 cycles per wave instruction against
// ~6 for ds_add_u64 on distinct cells), so scatters accumulate into a fixed-point LDS box
// of 64-bit cells with integer adds: every contribution c is stored as round(c * 2^s).
// Per channel and tile, s = 50 - e where 2^e bounds the largest |value|: every scaled
// contribution is below 2^50 (tap weights are at most 1) and no cell can overflow (at
// most 2^10 samples per tile touch a cell). Integer sums are exact and order-independent;
// each cell is rounded to float once, when it is flushed. A contribution errs by at most
// half the quantum 2^(e-50), i.e. 2^-50 of the tile's largest |value|, so a cell that
// receives n contributions from the tile carries at most n 2^-50 of it: far below the
// float rounding of the partial sums the flush adds into the image for every cell within
// about 2^-20 of the tile's largest value (relative precision, as with the reference's
// float atomics), an absolute 2^-44 of that value (n = 64) for the cells below.

// A channel's conversion factor mul: contribution c -> round(c * mul); cell q -> q / mul.
__device__ __forceinline__ double fixed_scale(float vmax) {
  int e = 50;                               // (vmax = 0: scale 0)
  if (vmax > 0.0f) (void)frexpf(vmax, &e);  // vmax < 2^e
  return ldexp(1.0, 50 - e);                // a power of two: products stay exact
}

// round(a * b) as int64 for |a * b| < 2^51: the FMA with 2^52 + 2^51 rounds the product
// to an integer held in the low mantissa bits (one FMA instead of a float -> int64
// conversion sequence; the low word of the constant is zero, so removing it costs one
// subtraction on the high word); fix2 adds two such products.
constexpr double kMagic = 6755399441055744.0;  // 2^52 + 2^51

__device__ __forceinline__ acc_t fix1(double a, double b) {
  return (acc_t)(__double_as_longlong(fma(a, b, kMagic)) - __double_as_longlong(kMagic));
}
__device__ __forceinline__ acc_t fix2(double a, double b, double c, double d) {
  return (acc_t)(__double_as_longlong(fma(a, b, fma(c, d, kMagic))) - __double_as_longlong(kMagic));
}

// q * inv rounded to float. The int64 -> double conversion (high word scaled by 2^32 plus
// the low word, one rounding) is exact below 2^53 in magnitude and takes no branch.
__device__ __forceinline__ float fixed_to_float(long long q, double inv) {
  return (float)((double)q * inv);
}

template <class T>
__device__ __forceinline__ void zero_acc(T* box, int cells) {
  for (int i = threadIdx.x; i < cells; i += kTB) box[i] = 0;
}

// Largest value over the workgroup (v >= 0); uses and releases wsum.
__device__ __forceinline__ float block_max(float v, float* wsum) {
  v = __ockl_wfred_max_f32(v);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  float m = wsum[0];
  for (int w = 1; w < kTB / 64; ++w) m = fmaxf(m, wsum[w]);
  __syncthreads();
  return m;
}

// Flush the fixed-point box into image channel `dst`: one folded global float atomic per
// non-zero cell. FLIPS as for staging.
template <bool FLIPS>
__device__ __forceinline__ void flush_rows(const acc_t* box, const TileState& S, double mul, float* __restrict__ dst,
                                           unsigned signs, RowWalk w) {
  const RowEnt* rt = reinterpret_cast<const RowEnt*>(pp_box + S.tab);
  const int rows = S.rows;
  const double inv = 1.0 / mul;  // (exact: mul is a power of two)
  constexpr int kFB = 4;  // cells read before the first is used (their LDS latencies overlap)
  for (; w.r < rows; w.r += kFB * w.rp) {
    RowEnt e[kFB];
    long long q[kFB];
    // table entries and cells read unconditionally, as in staging (rows past the end read
    // the last row's, a valid cell, and are masked below)
#pragma unroll
    for (int u = 0; u < kFB; ++u) e[u] = rt[min(w.r + u * w.rp, rows - 1)];
#pragma unroll
    for (int u = 0; u < kFB; ++u) q[u] = (long long)box[(e[u].cell & kCellMask) + w.c];
#pragma unroll
    for (int u = 0; u < kFB; ++u)
      if (w.r + u * w.rp < rows && q[u] != 0) {
        const float v = fixed_to_float(q[u], inv);
        unsafeAtomicAdd(dst + e[u].goff + w.fx, (FLIPS && cell_neg(e[u].cell, w, signs)) ? -v : v);
      }
  }
}

__device__ __forceinline__ void flush_acc(const acc_t* box, const TileState& S, double mul, float* __restrict__ dst,
                                          const KArgs& a, unsigned signs) {
  RowWalk w = row_walk(S, a);
  if (!w.active) return;
  if (a.bx == 2 || a.by == 2 || a.bz == 2) flush_rows<true>(box, S, mul, dst, signs, w);
  else flush_rows<false>(box, S, mul, dst, signs, w);
}

// ---- anisotropic degrees in the window of the largest -------------------------
// An axis of degree dp - 1 < DP - 1 has its dp taps at knots od .. od + dp - 1, od =
// ceil(x - dp/2), inside the DP-tap window o .. o + DP - 1 of o = ceil(x - DP/2): the shift
// sh = od - o lies in [0, DP - dp] (x - dp/2 >= x - DP/2, and od - o < (DP - dp)/2 + 1).
// The window's weights are the dp-tap weights (weights_ct<dp>, the same bits as the kernels
// compiled for that degree) at positions sh .. sh + dp - 1 and 0 elsewhere; placed with
// selects, not register indexing.
template <int DP, int D>
__device__ __forceinline__ void place_win(float x, int o, float* w, float* dw, float* hw) {
  const int od = (int)ceilf(x - 0.5f * (float)D);
  float a[D], b[D], c[D];
  weights_ct<D>(x - (float)od, a, b, c);
  const int sh = od - o;
#pragma unroll
  for (int i = 0; i < DP; ++i) {
    float wi = 0.f, di = 0.f, hi = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (i - k >= 0 && i - k <= DP - D && sh == i - k) {
        wi = a[k];
        di = b[k];
        hi = c[k];
      }
    w[i] = wi;
    dw[i] = di;
    hw[i] = hi;
  }
}

template <int DP>
__device__ __forceinline__ void weights_win(int dp, float x, int o, float* w, float* dw, float* hw) {
  if (dp >= DP) weights_ct<DP>(x - (float)o, w, dw, hw);
  else if (dp == 1) place_win<DP, 1>(x, o, w, dw, hw);
  else if (DP > 2 && dp == 2) place_win<DP, (DP > 2 ? 2 : 1)>(x, o, w, dw, hw);
  else if (DP > 3 && dp == 3) place_win<DP, (DP > 3 ? 3 : 1)>(x, o, w, dw, hw);
  else if (DP > 4) place_win<DP, (DP > 4 ? 4 : 1)>(x, o, w, dw, hw);
}

// The three axes of a sample on the global path (outside the box) and f(ax, ay, az) on
// them: DP taps each, or, with anisotropic degrees, each axis's own taps only (make_axis<0>,
// the runtime-degree forms of the per-sample kernels), so that no tap outside an axis's
// support is read (0 times a non-finite value would be NaN where the reference, which never
// reads that tap, has a value). cx, cy, cz: the axes' boundary codes.
template <int DP, int NEED, bool AN, class F>
__device__ __forceinline__ auto with_axes(const KArgs& a, float x, float y, float z, int cx, int cy, int cz, F f) {
  if constexpr (AN) {
    Axis<0> ax, ay, az;
    make_axis<0, NEED>(x, a.dpx, a.d0x, 1, cx, ax);
    make_axis<0, NEED>(y, a.dpy, a.d0y, a.d0x, cy, ay);
    make_axis<0, NEED>(z, a.dpz, a.d0z, a.d0x * a.d0y, cz, az);
    return f(ax, ay, az);
  } else {
    Axis<DP> ax, ay, az;
    make_axis<DP, NEED>(x, DP, a.d0x, 1, cx, ax);
    make_axis<DP, NEED>(y, DP, a.d0y, a.d0x, cy, ay);
    make_axis<DP, NEED>(z, DP, a.d0z, a.d0x * a.d0y, cz, az);
    return f(ax, ay, az);
  }
}
#define PP_AXIS_DP(ax) std::decay_t<decltype(ax)>::kDP

// Unsigned, unfolded weights of one in-window sample and its box offset
template <int DP>
struct LSample {
  float wx[DP], wy[DP], wz[DP], dx[DP], dy[DP], dz[DP];
  int base;
};

template <int DP, bool DERIV, bool AN>
__device__ __forceinline__ void lsample(const KArgs& a, float x, float y, float z, const TileState& S, int P, int Q,
                                        LSample<DP>& L) {
  const int ox = knot_fast<DP>(x), oy = knot_fast<DP>(y), oz = knot_fast<DP>(z);
  float h[DP];
  if constexpr (AN) {
    weights_win<DP>(a.dpx, x, ox, L.wx, L.dx, h);
    weights_win<DP>(a.dpy, y, oy, L.wy, L.dy, h);
    weights_win<DP>(a.dpz, z, oz, L.wz, L.dz, h);
  } else {
    weights_ct<DP>(x - (float)ox, L.wx, L.dx, h);
    weights_ct<DP>(y - (float)oy, L.wy, L.dy, h);
    weights_ct<DP>(z - (float)oz, L.wz, L.dz, h);
  }
  L.base = (ox - S.lo[0]) + P * (oy - S.lo[1]) + Q * (oz - S.lo[2]);
}

// ... with the second-derivative weights too (pull_hess and its pullback)
template <int DP>
struct LSampleH : LSample<DP> {
  float hx[DP], hy[DP], hz[DP];
};

template <int DP, bool AN>
__device__ __forceinline__ void lsample_h(const KArgs& a, float x, float y, float z, const TileState& S, int P, int Q,
                                          LSampleH<DP>& L) {
  const int ox = knot_fast<DP>(x), oy = knot_fast<DP>(y), oz = knot_fast<DP>(z);
  if constexpr (AN) {
    weights_win<DP>(a.dpx, x, ox, L.wx, L.dx, L.hx);
    weights_win<DP>(a.dpy, y, oy, L.wy, L.dy, L.hy);
    weights_win<DP>(a.dpz, z, oz, L.wz, L.dz, L.hz);
  } else {
    weights_ct<DP>(x - (float)ox, L.wx, L.dx, L.hx);
    weights_ct<DP>(y - (float)oy, L.wy, L.dy, L.hy);
    weights_ct<DP>(z - (float)oz, L.wz, L.dz, L.hz);
  }
  L.base = (ox - S.lo[0]) + P * (oy - S.lo[1]) + Q * (oz - S.lo[2]);
}

// ---- scatter slot order ------------------------------------------------------------
// LDS atomics of one instruction that hit the same cell serialise: on gfx950 a single
// such collision replays the instruction (two to three times the cost of one whose 64
// lanes hit distinct cells; banks matter little for these atomics). Samples that
// neighbour along x or z -- diagonals included -- often share a knot, and in tile_pos
// order they sit in neighbouring lanes of one slot. The scatter passes therefore visit a
// lane's slots rotated by the colour r = (x mod 2) + 2 (z mod 2) of its tile position: in
// one pass, lanes whose samples neighbour along x or z (or both) hold samples on different
// y rows, and lanes of one colour are two positions apart, where equal knots would need a
// twofold compression of the deformation.
template <int TX, int TY, int TZ>
__device__ __forceinline__ int scatter_rot() {
  constexpr int VPT = Shape<TX, TY, TZ>::VPT;
  static_assert(VPT >= 4 && (VPT & (VPT - 1)) == 0, "slots per thread: a power of two >= 4");
  const int t = threadIdx.x;
  return ((t % TX) & 1) + 2 * (((t / TX) % TZ) & 1);
}

// Keeps the scheduler from interleaving the unrolled per-sample bodies (each body is
// long enough to hide its own latencies among the resident waves; interleaving them
// only multiplies the live registers).
__device__ __forceinline__ void sample_fence() { __builtin_amdgcn_sched_barrier(0); }

// a[q] <- a[(q + r) mod N] for a lane-dependent r (selects, no register indexing).
template <int N, class T>
__device__ __forceinline__ void rotate_by(T (&a)[N], int r) {
#pragma unroll
  for (int s = 1; s < N; s <<= 1) {
    T b[N];
#pragma unroll
    for (int q = 0; q < N; ++q) b[q] = (r & s) ? a[(q + s) % N] : a[q];
#pragma unroll
    for (int q = 0; q < N; ++q) a[q] = b[q];
  }
}

// Bit q of the result is bit (q + r) mod N of m.
template <int N>
__device__ __forceinline__ unsigned rotate_bits(unsigned m, int r) {
  constexpr unsigned mask = (1u << N) - 1u;
  m &= mask;
  return ((m >> r) | (m << ((N - r) & (N - 1)))) & mask;
}

// Fixed-point contribution of tap weight w times the row factor c (scaled) to a 64-bit cell.
__device__ __forceinline__ void lds_add_cell(acc_t* cell, float w, double c) { atomicAdd(cell, fix1(w, c)); }

// Scatter one in-window sample's value (s0 = value * scale) into the fixed-point box.
template <int DP>
__device__ __forceinline__ void lds_scatter_val(acc_t* box, const LSample<DP>& L, int P, int Q, double s0) {
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    acc_t* plane = box + L.base + Q * k;
    const double sk = (double)L.wz[k] * s0;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      const double wv = (double)L.wy[j] * sk;
#pragma unroll
      for (int i = 0; i < DP; ++i) lds_add_cell(plane + P * j + i, L.wx[i], wv);
    }
  }
}

// ---- push / push_grad -------------------------------------------------------------
// One tile of push (GRAD=false) or push_grad (GRAD=true) for batch index bb. The caller
// provides the LDS box (cap floats, 8-byte aligned; used as 64-bit cells), wsum (kTB/64
// floats) and the tile state.
template <int DP, int TX, int TY, int TZ, bool GRAD, bool AN = false>
__device__ __forceinline__ void push_tile(const KArgs& a, int tile, int bb, int cap, acc_t* box, float* wsum,
                                          TileState& S) {
  constexpr int VPT = Shape<TX, TY, TZ>::VPT;
  constexpr int NG = GRAD ? 3 : 1;
  const float* phi = a.phi + bb * a.phi_bs;
  const long long n1 = a.n1;
  // the first channel's values are loaded before the deformation is read
  float g[NG][VPT];
  const float* src0 = a.f + bb * a.f_bs;
  const bool read0 = GRAD || (a.vmode == 0 && a.nchan > 0);
  auto prefetch = [&](const Vox<VPT>& W) {  // (unconditional loads, masked after: see tile_setup)
    float t[NG][VPT];
    if (read0) {
#pragma unroll
      for (int q = 0; q < VPT; ++q)
#pragma unroll
        for (int e = 0; e < NG; ++e) t[e][q] = src0[W.v[q] + e * n1];
    } else {
#pragma unroll
      for (int q = 0; q < VPT; ++q)
#pragma unroll
        for (int e = 0; e < NG; ++e) t[e][q] = 0.f;
    }
#pragma unroll
    for (int q = 0; q < VPT; ++q)
#pragma unroll
      for (int e = 0; e < NG; ++e) g[e][q] = ((W.inside >> q) & 1u) ? t[e][q] : 0.f;
  };
  // channel 0's scale bound: the largest finite per-sample bound over all samples in the
  // grid (a superset of the in-window ones, so a valid bound), reduced with the box
  auto bound0 = [&]() {
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      float b = fabsf(g[0][q]);
      if (GRAD) b += fabsf(g[NG > 1 ? 1 : 0][q]) + fabsf(g[NG > 2 ? 2 : 0][q]);
      if (isfinite(b)) m = fmaxf(m, b);
    }
    return m;
  };
  Vox<VPT> V;
  float bound_c0 = 0.f;
  tile_setup<DP, TX, TY, TZ, false, true>(a, phi, tile, cap, V, S, prefetch, bound0, box, &bound_c0);
  const bool lds = S.lds != 0;
  const int P = S.Pc, Q = S.Qc;
  const int srot = scatter_rot<TX, TY, TZ>();
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
        int px, py, pz;
        tile_pos<TX, TY, TZ>(l, px, py, pz);
        const int id = ax == 0 ? S.org[0] + px : ax == 1 ? S.org[1] + py : S.org[2] + pz;
#pragma unroll
        for (int e = 0; e < NG; ++e) {
          const float t = vmode == 0 ? src[V.v[q] + e * n1] : vmode == 1 ? 1.0f : (float)(id + 1);
          g[e][q] = ((V.valid >> q) & 1u) ? t : 0.f;
        }
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
        if (isfinite(b)) vmax = fmaxf(vmax, b);
        else use &= ~(1u << q);
      }
      if (c == 0 && read0) {  // bound reduced and box zeroed within tile_setup's barriers
        scale = fixed_scale(bound_c0);
      } else {
        const float m = block_max(vmax, wsum);
        scale = fixed_scale(m);
        zero_acc(box, S.cells_c);
        __syncthreads();
      }
    }
    if (lds) {
      // the box, one pass per slot in the lane's order (scatter_rot)
      float sx[VPT], sy[VPT], sz[VPT], sg[NG][VPT];
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        sx[q] = V.x[q];
        sy[q] = V.y[q];
        sz[q] = V.z[q];
#pragma unroll
        for (int e = 0; e < NG; ++e) sg[e][q] = g[e][q];
      }
      rotate_by(sx, srot);
      rotate_by(sy, srot);
      rotate_by(sz, srot);
#pragma unroll
      for (int e = 0; e < NG; ++e) rotate_by(sg[e], srot);
      const unsigned suse = rotate_bits<VPT>(use, srot);
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        if (!((suse >> q) & 1u)) continue;
        LSample<DP> L;
        lsample<DP, GRAD, AN>(a, sx[q], sy[q], sz[q], S, P, Q, L);
        if (GRAD) {
          const double s0 = (double)sg[0][q] * scale;
          const double s1 = (double)sg[NG > 1 ? 1 : 0][q] * scale;
          const double s2 = (double)sg[NG > 2 ? 2 : 0][q] * scale;
#pragma unroll
          for (int k = 0; k < DP; ++k)
#pragma unroll
            for (int j = 0; j < DP; ++j) {
              acc_t* row = box + L.base + P * j + Q * k;
              const double cx = (double)(L.wz[k] * L.wy[j]) * s0;
              const double cyz = (double)(L.wz[k] * L.dy[j]) * s1 + (double)(L.dz[k] * L.wy[j]) * s2;
#pragma unroll
              for (int i = 0; i < DP; ++i) atomicAdd(row + i, fix2(L.dx[i], cx, L.wx[i], cyz));
            }
        } else {
          lds_scatter_val<DP>(box, L, P, Q, (double)sg[0][q] * scale);
        }
        sample_fence();
      }
    }
    // samples outside the box (rare): through global memory, per sample
    unsigned glob = V.valid & ~use;
    if (!GRAD) {
#pragma unroll
      for (int q = 0; q < VPT; ++q)
        if (!isfinite(g[0][q])) glob &= ~(1u << q);  // push skips non-finite values (pushpull.c:114)
    }
    if (__any(glob != 0u)) {  // one slot at a time, rotating the arrays (VPT rotations restore them)
#pragma unroll 1
      for (int q = 0; q < VPT; ++q, V.rotate(), rot_rows(g)) {
        if (!((glob >> q) & 1u)) continue;
        constexpr int NEED = GRAD ? (kW | kDW) : kW;
        with_axes<DP, NEED, AN>(a, V.x[0], V.y[0], V.z[0], chan_code(a, c, 0, a.bx), chan_code(a, c, 1, a.by),
                            chan_code(a, c, 2, a.bz), [&](const auto& ax, const auto& ay, const auto& az) {
                              constexpr int D = PP_AXIS_DP(ax);
                              if (GRAD) scatter_grad<D>(dst, ax, ay, az, g[0][0], g[NG > 1 ? 1 : 0][0], g[NG > 2 ? 2 : 0][0]);
                              else scatter_val<D>(dst, ax, ay, az, g[0][0]);
                              return 0;
                            });
      }
    }
    if (lds) {
      __syncthreads();
      flush_acc(box, S, scale, dst, a, chan_signs(a, c));
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
  return DP == 2 ? (KIND == 7 ? 5 : 0)
       : DP == 3 ? (KIND == 0 || KIND == 4 ? 6 : KIND == 1 || KIND == 5 || KIND == 7 ? 5 : KIND == 2 || KIND == 3 ? 4 : 0)
       : DP == 4 ? (KIND == 2 || KIND == 3 || KIND == 6 ? 4 : KIND == 7 || KIND == 0 ? 5 : 0)
                 : 0;
}

// AN: anisotropic degrees (each axis in the window of DP taps; compiler-chosen occupancy).
template <int DP, int TX, int TY, int TZ, bool GRAD, bool AN = false>
__global__ __launch_bounds__(kTB, (AN ? 0 : tile_waves<DP, GRAD ? 8 : 7>())) void k_push_tile(KArgs a, int cap) {
  __shared__ TileState S;
  __shared__ float wsum[kTB / 64];
  push_tile<DP, TX, TY, TZ, GRAD, AN>(a, xcd_block(blockIdx.x, gridDim.x), blockIdx.y, cap,
                                  reinterpret_cast<acc_t*>(pp_box), wsum, S);
}

// ---- gathers: pull (KIND 0), phi pullback of pull (1), fused objective (2), the ------
// ---- vector-Jacobian product of pull (3: push(delta) and the phi pullback together), -
// ---- pull_grad (4), pull_hess (5) and the pullback of pull_grad (6) ------------------
// The contractions run one axis at a time -- x within each row, then y within each
// plane, then z -- which takes fewer FMAs than weighting every row by its yz product
// (a cubic gradient: 188 instead of 224 per sample).
// The dp x dp taps of box plane k of a sample, loaded together: the reads of a plane are
// all issued before the first is used (left to itself the scheduler issues one read and
// waits for it before the next, which leaves every read's latency exposed).
template <int DP>
__device__ __forceinline__ void lds_plane(const float* box, const LSample<DP>& L, int P, int Q, int k,
                                          float (&v)[DP][DP]) {
  const float* pl = box + L.base + Q * k;
#pragma unroll
  for (int j = 0; j < DP; ++j)
#pragma unroll
    for (int i = 0; i < DP; ++i) v[j][i] = pl[P * j + i];
  sample_fence();
}

template <int DP>
__device__ __forceinline__ float lds_val(const float* box, const LSample<DP>& L, int P, int Q) {
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    float v[DP][DP];
    lds_plane<DP>(box, L, P, Q, k, v);
    float p = 0.f;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      float r = 0.f;
#pragma unroll
      for (int i = 0; i < DP; ++i) r += L.wx[i] * v[j][i];
      p += L.wy[j] * r;
    }
    acc += L.wz[k] * p;
  }
  return acc;
}

template <int DP, bool VAL>
__device__ __forceinline__ void lds_grad(const float* box, const LSample<DP>& L, int P, int Q, float& val,
                                         float& gx, float& gy, float& gz) {
  val = gx = gy = gz = 0.f;
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    float v[DP][DP];
    lds_plane<DP>(box, L, P, Q, k, v);
    float pv = 0.f, px = 0.f, py = 0.f;  // plane k: sum_j wy r, sum_j wy rd, sum_j dy r
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      float r = 0.f, rd = 0.f;
#pragma unroll
      for (int i = 0; i < DP; ++i) {
        r += L.wx[i] * v[j][i];
        rd += L.dx[i] * v[j][i];
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
__device__ __forceinline__ void lds_hess(const float* box, const LSampleH<DP>& L, int P, int Q, float (&h)[6]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) h[q] = 0.f;
#pragma unroll
  for (int k = 0; k < DP; ++k) {
    float v[DP][DP];
    lds_plane<DP>(box, L, P, Q, k, v);
    float pw = 0.f, pwd = 0.f, pwh = 0.f, pd = 0.f, pdd = 0.f, ph = 0.f;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
      float r = 0.f, rd = 0.f, rh = 0.f;
#pragma unroll
      for (int i = 0; i < DP; ++i) {
        r += L.wx[i] * v[j][i];
        rd += L.dx[i] * v[j][i];
        rh += L.hx[i] * v[j][i];
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

// Gradient gather of one sample of channel c outside the tile's box, through global
// memory; for the objective (KIND 2) it also scatters the residual, for the VJP (3) the
// upstream gradient. Returns the weight w of the pullback (the upstream gradient, or the
// residual). Rarely taken, so it gathers lean: it must not set the register budget of the
// whole kernel.
template <int DP, int KIND, bool AN>
__device__ __forceinline__ float outlier_grad(const KArgs& a, int c, const float* __restrict__ src,
                                              float* __restrict__ dst2, float x, float y, float z, float aux, float& gx,
                                              float& gy, float& gz) {
  return with_axes<DP, kW | kDW, AN>(a, x, y, z, chan_code(a, c, 0, a.bx), chan_code(a, c, 1, a.by),
                                     chan_code(a, c, 2, a.bz), [&](const auto& ax, const auto& ay, const auto& az) {
    constexpr int D = PP_AXIS_DP(ax);
    float val = 0.f;
    if (KIND == 2) sample_val_grad<D, true>(src, ax, ay, az, val, gx, gy, gz);
    else sample_grad<D, true>(src, ax, ay, az, gx, gy, gz);
    const float w = (KIND == 2) ? val - aux : aux;
    if ((KIND == 2 || KIND == 3) && isfinite(w)) scatter_val<D>(dst2, ax, ay, az, w);
    return w;
  });
}


// One tile of pull (KIND 0), the pull pullback (1), the fused objective (2), the VJP
// of pull (3: dphi = sum_c delta_c grad f_c and df0 += push(delta) from one staging),
// pull_grad (4: the three components of grad f_c per channel), pull_hess (5: the 3x3
// Hessian per channel) or the pullback of pull_grad (6: dphi_b = sum_{a,c} delta_{a,c} H_{a,b,c}).
// box: cap floats (8-byte aligned; the scattering kinds reuse it for the accumulator);
// wsum: kTB/64 floats.
// The pullback of every sample is summed over channels in registers (gsum), and the
// objective keeps each sample's residual in registers between its two passes.
template <int DP, int TX, int TY, int TZ, int KIND, bool AN = false>
__device__ __forceinline__ void gather_tile(const KArgs& a, int tile, int bb, int cap, float* box, float* wsum,
                                            TileState& S) {
  constexpr int VPT = Shape<TX, TY, TZ>::VPT;
  constexpr bool AUX = KIND >= 1 && KIND <= 3;  // reads an upstream gradient or target per sample
  constexpr bool RED = AUX || KIND == 6;         // writes dphi, reduced over the channels
  const float* phi = a.phi + bb * a.phi_bs;
  float auxv[VPT];  // upstream gradient / target; channel 0's is loaded before phi is read
  auto prefetch = [&](const Vox<VPT>& W) {  // (unconditional loads, masked after: see tile_setup)
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      const float t = AUX ? a.aux[bb * a.aux_bs + W.v[q]] : 0.f;
      auxv[q] = ((W.inside >> q) & 1u) ? t : 0.f;
    }
  };
  Vox<VPT> V;
  tile_setup<DP, TX, TY, TZ, true, KIND == 2 || KIND == 3>(a, phi, tile, cap, V, S, prefetch);
  const bool lds = S.lds != 0;
  const int P = S.Pf, Q = S.Qf;
  const long long n1 = a.n1;
  float gsum[3][VPT];
#pragma unroll
  for (int q = 0; q < VPT; ++q) gsum[0][q] = gsum[1][q] = gsum[2][q] = 0.f;
  float lsum = 0.f;
  const float nan = __builtin_nanf("");
  // samples read from the box, and valid samples taking the global path (outside the
  // box's window, or the whole tile when its box exceeds the budget)
  const unsigned lmask = lds ? (V.valid & V.inl) : 0u;
  const unsigned omask = V.valid & ~lmask;
  const bool any_outlier = __any(omask != 0u);
  for (int c = 0; c < a.nchan; ++c) {
    const float* src = a.f + bb * a.f_bs + c * a.f_cs;
    // KIND 2/3: in-window samples with a finite weight, scattered via LDS; their weight
    // (the residual, which replaces the target in auxv, or the upstream gradient) is auxv
    unsigned rfin = 0u;
    if (AUX && c > 0) {  // loaded before staging so both overlap (unconditional loads, masked after)
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        const float t = a.aux[bb * a.aux_bs + c * a.aux_cs + V.v[q]];
        auxv[q] = ((V.valid >> q) & 1u) ? t : 0.f;
      }
    }
    unsigned lm = lmask, om = omask;  // this channel's box and global-path samples
    bool anyo = any_outlier;
    if (lds) {
      stage_box(box, S, src, a, chan_signs(a, c));
      __syncthreads();
      if constexpr (AN) {
        const bool bad = sanitize_box(box, S);
        if (__any(bad) && (threadIdx.x & 63) == 0) atomicMax(&S.nonfin, c + 1);
        __syncthreads();
        if (uni(S.nonfin) == c + 1) {
          lm = 0u;
          om = V.valid;
          anyo = true;
        }
      }
    }
    // invalid samples: NaN outputs
    if (KIND == 0 || KIND == 4 || KIND == 5) {
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        if (!((V.inside >> q) & 1u) || ((V.valid >> q) & 1u)) continue;
        float* o = a.out + bb * a.out_bs + c * a.out_cs + V.v[q];
        if (KIND == 0) o[0] = nan;
        if (KIND == 4) put3(o, n1, nan, nan, nan, 0);
        if (KIND == 5)
          for (int e = 0; e < 9; ++e) o[e * n1] = nan;
      }
    }
    // One valid sample: array slot qi (its slot q, or 0 in the rotating outlier loop),
    // from the box (use_lds) or through global memory.
    auto body = [&](const int qi, const int q, const bool use_lds) {
      const int v = V.v[qi];
      const float x = V.x[qi], y = V.y[qi], z = V.z[qi];
      if (KIND == 0) {
        float val;
        if (use_lds) {
          LSample<DP> L;
          lsample<DP, false, AN>(a, x, y, z, S, P, Q, L);
          val = lds_val<DP>(box, L, P, Q);
        } else {
          val = with_axes<DP, kW, AN>(a, x, y, z, chan_code(a, c, 0, a.bx), chan_code(a, c, 1, a.by),
                                  chan_code(a, c, 2, a.bz), [&](const auto& ax, const auto& ay, const auto& az) {
                                    return sample_val<PP_AXIS_DP(ax), true>(src, ax, ay, az);
                                  });
        }
        if (a.vscale != 1.0f || a.add_id) {
          int px, py, pz;
          tile_pos<TX, TY, TZ>(threadIdx.x + kTB * q, px, py, pz);
          val = pull_out(a, val, c, S.org[0] + px, S.org[1] + py, S.org[2] + pz);
        }
        a.out[bb * a.out_bs + c * a.out_cs + v] = val;
        return;
      }
      if (KIND == 5 || KIND == 6) {
        float h[6];
        if (use_lds) {
          LSampleH<DP> L;
          lsample_h<DP, AN>(a, x, y, z, S, P, Q, L);
          lds_hess<DP>(box, L, P, Q, h);
        } else {
          with_axes<DP, kW | kDW | kHW, AN>(a, x, y, z, chan_code(a, c, 0, a.bx), chan_code(a, c, 1, a.by),
                                            chan_code(a, c, 2, a.bz), [&](const auto& ax, const auto& ay, const auto& az) {
                                          sample_hess<PP_AXIS_DP(ax)>(src, ax, ay, az, h);
                                          return 0;
                                        });
        }
        if (KIND == 5) {  // column-major 3x3 (pushpull.c:88-93)
          float* o = a.out + bb * a.out_bs + c * a.out_cs + v;
          o[0] = h[0]; o[4 * n1] = h[1]; o[8 * n1] = h[2];
          o[1 * n1] = h[3]; o[3 * n1] = h[3];
          o[2 * n1] = h[4]; o[6 * n1] = h[4];
          o[5 * n1] = h[5]; o[7 * n1] = h[5];
        } else {
          const float* dl = a.aux + bb * a.aux_bs + c * a.aux_cs + v;
          const float d0 = dl[0], d1 = dl[n1], d2 = dl[2 * n1];
          gsum[0][qi] += d0 * h[0] + d1 * h[3] + d2 * h[4];
          gsum[1][qi] += d0 * h[3] + d1 * h[1] + d2 * h[5];
          gsum[2][qi] += d0 * h[4] + d1 * h[5] + d2 * h[2];
        }
        return;
      }
      if (KIND == 4) {
        float gx, gy, gz;
        if (use_lds) {
          LSample<DP> L;
          lsample<DP, true, AN>(a, x, y, z, S, P, Q, L);
          float val;
          lds_grad<DP, false>(box, L, P, Q, val, gx, gy, gz);
        } else {
          with_axes<DP, kW | kDW, AN>(a, x, y, z, chan_code(a, c, 0, a.bx), chan_code(a, c, 1, a.by),
                                      chan_code(a, c, 2, a.bz), [&](const auto& ax, const auto& ay, const auto& az) {
            sample_grad<PP_AXIS_DP(ax), true>(src, ax, ay, az, gx, gy, gz);
            return 0;
          });
        }
        put3(a.out + bb * a.out_bs + c * a.out_cs + v, n1, gx, gy, gz, 0);
        return;
      }
      const float aux = auxv[qi];
      float val = 0.f, gx, gy, gz, w;
      if (use_lds) {
        LSample<DP> L;
        lsample<DP, true, AN>(a, x, y, z, S, P, Q, L);
        lds_grad<DP, KIND == 2>(box, L, P, Q, val, gx, gy, gz);
        w = (KIND == 2) ? val - aux : aux;
        if ((KIND == 2 || KIND == 3) && isfinite(w)) {
          if (KIND == 2) auxv[qi] = w;
          rfin |= 1u << q;
        }
      } else {
        w = outlier_grad<DP, KIND, AN>(a, c, src, a.out2 + bb * a.out2_bs + c * a.out2_cs, x, y, z, aux, gx, gy, gz);
      }
      if (KIND == 2 && isfinite(w)) lsum += 0.5f * w * w;
      gsum[0][qi] += w * gx;
      gsum[1][qi] += w * gy;
      gsum[2][qi] += w * gz;
    };
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      if ((lm >> q) & 1u) body(q, q, true);
      sample_fence();
    }
    if (anyo) {  // rare: one slot at a time, rotating the arrays (VPT rotations restore them)
#pragma unroll 1
      for (int q = 0; q < VPT; ++q, V.rotate(), rot(auxv), rot_rows(gsum))
        if ((om >> q) & 1u) body(0, q, false);
    }
    if (lds) {
      __syncthreads();
      if (KIND == 2 || KIND == 3) {  // scatter the residuals (gradients) of the in-window samples via the fixed-point box
        float vb = 0.f;
#pragma unroll
        for (int q = 0; q < VPT; ++q)
          if ((rfin >> q) & 1u) vb = fmaxf(vb, fabsf(auxv[q]));
        const float m = block_max(vb, wsum);  // also orders the staging reads
        const double scale = fixed_scale(m);
        acc_t* cbox = reinterpret_cast<acc_t*>(box);
        zero_acc(cbox, S.cells_c);
        __syncthreads();
        // slots in the lane's rotated order (scatter_rot)
        const int srot = scatter_rot<TX, TY, TZ>();
        float sx[VPT], sy[VPT], sz[VPT], sw[VPT];
#pragma unroll
        for (int q = 0; q < VPT; ++q) {
          sx[q] = V.x[q];
          sy[q] = V.y[q];
          sz[q] = V.z[q];
          sw[q] = auxv[q];
        }
        rotate_by(sx, srot);
        rotate_by(sy, srot);
        rotate_by(sz, srot);
        rotate_by(sw, srot);
        const unsigned sfin = rotate_bits<VPT>(rfin, srot);
#pragma unroll
        for (int q = 0; q < VPT; ++q) {
          if (!((sfin >> q) & 1u)) continue;
          LSample<DP> L;
          lsample<DP, false, AN>(a, sx[q], sy[q], sz[q], S, S.Pc, S.Qc, L);
          lds_scatter_val<DP>(cbox, L, S.Pc, S.Qc, (double)sw[q] * scale);
          sample_fence();
        }
        __syncthreads();
        flush_acc(cbox, S, scale, a.out2 + bb * a.out2_bs + c * a.out2_cs, a, chan_signs(a, c));
        __syncthreads();
      }
    }
  }
  if (RED) {
    const bool step = KIND == 2 && a.phi_next;
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      if (!((V.inside >> q) & 1u)) continue;
      const int v = V.v[q];
      const bool ok = (V.valid >> q) & 1u;
      float* o = a.out + bb * a.out_bs + v;
      float g0 = ok ? gsum[0][q] : nan, g1 = ok ? gsum[1][q] : nan, g2 = ok ? gsum[2][q] : nan;
      if (ok && a.accum) {
        g0 += o[0];
        g1 += o[n1];
        g2 += o[2 * n1];
      }
      o[0] = g0;
      o[n1] = g1;
      o[2 * n1] = g2;
      if (step) {  // phi + step * dL/dphi, as pp_axpy rounds it
        float* pn = a.phi_next + bb * a.phi_bs + v;
        const float p0 = phi[v], p1 = phi[v + n1], p2 = phi[v + 2 * n1];
        pn[0] = fmaf(a.step, g0, p0);
        pn[n1] = fmaf(a.step, g1, p1);
        pn[2 * n1] = fmaf(a.step, g2, p2);
      }
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

template <int DP, int TX, int TY, int TZ, int KIND, bool AN = false>
__global__ __launch_bounds__(kTB, (AN ? 0 : tile_waves<DP, KIND>())) void k_gather_tile(KArgs a, int cap) {
  __shared__ TileState S;
  __shared__ float wsum[kTB / 64];
  gather_tile<DP, TX, TY, TZ, KIND, AN>(a, xcd_block(blockIdx.x, gridDim.x), blockIdx.y, cap, pp_box, wsum, S);
}

// ---- batched launch over volumes of uneven size ----------------------------------------
// A plan orders its volumes by degree; each degree present (the largest, for a volume with
// a degree per axis, which gets kernels of its own: AN) gets one launch of a kernel compiled
// for that degree only (a single kernel switching over five degrees would carry the
// register footprint of all of them). Within a launch the tiles of the group's volumes
// are numbered consecutively (BatchVol::tile0, relative to the group) and cut into eight
// contiguous ranges of equal work, one per XCD (workgroup b runs on XCD b % 8), so that
// neighbouring tiles share an L2 and no XCD idles while another still holds the heavy
// (many-channel) volumes. Degrees 1-4 use the LDS tile bodies; degree 0 runs the same
// bodies with an empty box budget, i.e. per sample through global memory.
constexpr int kBTX = 16, kBTY = 8, kBTZ = 8;

// k_batch kinds: gather kinds 0-2, push 3, push_grad 4, the pull VJP 5 (gather kind 3)
constexpr int batch_gather_kind(int kind) { return kind <= 2 ? kind : kind == 5 ? 3 : -1; }
constexpr int batch_waves_kind(int kind) { return kind <= 2 ? kind : kind == 5 ? 3 : kind + 4; }

template <int DP, int KIND, bool AN = false>
__global__ __launch_bounds__(kTB, (AN ? 0 : tile_waves<DP, batch_waves_kind(KIND)>())) void k_batch(
    const BatchVol* __restrict__ vols, int nvol, XcdSplit split, int cap) {
  constexpr int KG = batch_gather_kind(KIND) >= 0 ? batch_gather_kind(KIND) : 0;
  constexpr bool GATHER = batch_gather_kind(KIND) >= 0;
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
  const int c = DP >= 2 ? cap : 0;  // degree 0: no box
  if (GATHER) gather_tile<DP, kBTX, kBTY, kBTZ, KG, AN>(a, tile, 0, c, pp_box, wsum, S);
  else push_tile<DP, kBTX, kBTY, kBTZ, KIND == 4, AN>(a, tile, 0, c, reinterpret_cast<acc_t*>(pp_box), wsum, S);
}


// Batch kernel of op for degree dp (dp 1-5; AN: a degree per axis, dp the largest).
template <int KIND, bool AN>
static const void* batch_kernel(int dp) {
  switch (dp) {
    case 1:  // degree 0 on every axis (never a group with differing degrees)
      if constexpr (AN) return nullptr;
      else return reinterpret_cast<const void*>(&k_batch<1, KIND, AN>);
    case 2: return reinterpret_cast<const void*>(&k_batch<2, KIND, AN>);
    case 3: return reinterpret_cast<const void*>(&k_batch<3, KIND, AN>);
    case 4: return reinterpret_cast<const void*>(&k_batch<4, KIND, AN>);
    default: return reinterpret_cast<const void*>(&k_batch<5, KIND, AN>);
  }
}

template <bool AN>
static const void* batch_kernel_for(Op op, int dp) {
  switch (op) {
    case OP_PULL: return batch_kernel<0, AN>(dp);
    case OP_PULL_BWD: return batch_kernel<1, AN>(dp);
    case OP_OBJGRAD: return batch_kernel<2, AN>(dp);
    case OP_PULL_VJP: return batch_kernel<5, AN>(dp);
    case OP_PUSH: return batch_kernel<3, AN>(dp);
    case OP_PUSHG: return batch_kernel<4, AN>(dp);
    default: return nullptr;
  }
}

// ---- per-shape launch -------------------------------------------------------------
static size_t cap_bytes(int cap) { return (size_t)cap * sizeof(float); }

template <int TX, int TY, int TZ>
static int tiles_of(const KArgs& a) {
  return ((a.d1x + TX - 1) / TX) * ((a.d1y + TY - 1) / TY) * ((a.d1z + TZ - 1) / TZ);
}

template <int DP, int TX, int TY, int TZ, bool AN = false>
static hipError_t launch_shape(Op op, const KArgs& a, int nbatch, hipStream_t s, int cap) {
  const dim3 grid((unsigned)tiles_of<TX, TY, TZ>(a), (unsigned)nbatch, 1), block(kTB);
  const size_t sh = cap_bytes(cap);
  switch (op) {
    case OP_PULL: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 0, AN>), grid, block, sh, s, a, cap); break;
    case OP_PULLG: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 4, AN>), grid, block, sh, s, a, cap); break;
    case OP_PULLH: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 5, AN>), grid, block, sh, s, a, cap); break;
    case OP_PULLG_BWD: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 6, AN>), grid, block, sh, s, a, cap); break;
    case OP_PULL_BWD: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 1, AN>), grid, block, sh, s, a, cap); break;
    case OP_OBJGRAD: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 2, AN>), grid, block, sh, s, a, cap); break;
    case OP_PULL_VJP: hipLaunchKernelGGL((k_gather_tile<DP, TX, TY, TZ, 3, AN>), grid, block, sh, s, a, cap); break;
    case OP_PUSH: hipLaunchKernelGGL((k_push_tile<DP, TX, TY, TZ, false, AN>), grid, block, sh, s, a, cap); break;
    case OP_PUSHG: hipLaunchKernelGGL((k_push_tile<DP, TX, TY, TZ, true, AN>), grid, block, sh, s, a, cap); break;
    default: return hipErrorNotSupported;
  }
  return hipGetLastError();
}

}  // namespace pp
