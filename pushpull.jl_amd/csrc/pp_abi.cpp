// pp_abi.cpp -- extern "C" entry points of libpushpull_hip.so (declared in
// include/pushpull_hip.h): argument checking, stride set-up, Tier A host staging.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "pp_kernels.hpp"
#include "pp_memcheck.hpp"
#include "pp_sparse.hpp"
#include "pushpull_hip.h"

namespace {

thread_local char t_err[512];  // fixed storage: reporting an error allocates nothing
thread_local int t_status = 0;

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof t_err, fmt, ap);
  va_end(ap);
  t_status = code;
  return code;
}

int ok() {
  t_status = PP_OK;
  return PP_OK;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(PP_EDEVICE, "%s: %s", where, hipGetErrorString(e));
}

// The Tier B memory contract: every non-null array is device memory of the current device.
struct Arr {
  const void* p;
  const char* name;
};
int check_mem(std::initializer_list<Arr> arrs) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  for (const Arr& a : arrs) {
    if (!a.p) continue;
    if (const char* why = pp::device_memory_issue(a.p, dev))
      return fail(PP_EINVAL, "`%s` is %s; device arrays must be hipMalloc'ed memory of device %d", a.name, why, dev);
  }
  return PP_OK;
}

constexpr long long kMaxVox = (1LL << 31) - 1;  // per-volume voxel count (32-bit in-volume offsets)
constexpr long long kMaxDim = 1LL << 29;        // per-axis size (2*d must fit an int)
constexpr int kBlock = 256;
constexpr int kMaxCap = 16384;  // LDS box budget limit in floats (64 KiB of dynamic LDS)

// ---- kernel selection --------------------------------------------------------------
// Variants: 0 = one work-item per sample; 1 + k = tiled with tile shape k.
constexpr int kVariants = 4;
std::atomic<int> g_tiling{1};  // PP_OPT_TILING: 0 per sample, 1 automatic, 2 tiled
std::atomic<int> g_shape{0};   // PP_OPT_TILE_SHAPE (default when nothing is tuned)
std::atomic<int> g_tuned{1};   // PP_OPT_USE_TUNED
std::atomic<int> g_cap{0};     // PP_OPT_BOX_CAP: LDS box budget in floats (0: per-op default)

// A kernel choice: the variant and, for the tiled ones, the LDS box budget in floats
// (0: PP_OPT_BOX_CAP, else the op's default). The budget sets how many workgroups share a
// CU (LDS occupancy) against how much of a tile's footprint stays in LDS.
struct Choice {
  int variant;
  int cap;
};
thread_local Choice t_force{-1, 0};  // forced while pp_tune times candidates

// Autotuner table: (op, dp, log2 of the phi-grid size) -> choice.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, Choice> g_tune;

// Budgets pp_tune tries besides the op's default: the largest box that still lets 6, 5, 4
// and 3 workgroups share a CU's 160 KiB of LDS.
constexpr int kTuneCaps[] = {6770, 8136, 10184, 13596};
constexpr int kMaxCandidates = 1 + (kVariants - 1) * (1 + (int)(sizeof kTuneCaps / sizeof kTuneCaps[0]));
thread_local int t_ncand = 0;
thread_local Choice t_cand[kMaxCandidates];
thread_local float t_cand_ms[kMaxCandidates];

// Tuning-table key of a degree triple (dp = degree + 1 per axis): dp when isotropic, else
// one key per triple.
int dp_key(int dpx, int dpy, int dpz) { return (dpx == dpy && dpx == dpz) ? dpx : 100 + 36 * dpx + 6 * dpy + dpz; }

int size_class(long long n) {
  int c = 0;
  while ((1LL << (c + 1)) <= n) ++c;
  return c;
}

hipError_t launch_variant(pp::Op op, const pp::KArgs& a, int nbatch, hipStream_t s, int variant, int cap = 0) {
  if (variant > 0) {
    if (cap <= 0) cap = g_cap.load(std::memory_order_relaxed);
    hipError_t e = pp::launch_tiled(op, a, nbatch, s, pp::TileCfg{variant - 1, cap});
    if (e != hipErrorNotSupported) return e;
  }
  return pp::launch_op(op, a, nbatch, s, kBlock);
}

hipError_t launch(pp::Op op, const pp::KArgs& a, int nbatch, hipStream_t s) {
  if (t_force.variant >= 0) return launch_variant(op, a, nbatch, s, t_force.variant, t_force.cap);
  const int tiling = g_tiling.load(std::memory_order_relaxed);
  const int tiled = 1 + g_shape.load(std::memory_order_relaxed);
  if (tiling == 0) return launch_variant(op, a, nbatch, s, 0);
  if (tiling == 2) return launch_variant(op, a, nbatch, s, tiled);
  if (g_tuned.load(std::memory_order_relaxed)) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find({(int)op, dp_key(a.dpx, a.dpy, a.dpz), size_class(a.n1)});
    if (it != g_tune.end()) return launch_variant(op, a, nbatch, s, it->second.variant, it->second.cap);
  }
  // Built-in choice: the tiled kernels (at 256^3 the trilinear gathers run 10-25% faster
  // tiled than per sample, three channels with sliding boundaries twice as fast); a
  // trilinear gather over a flat list of samples (a one-row grid, as Tier A passes: a tile
  // would hold a single row of 16 samples) runs per sample.
  const bool trilinear = a.dpx == 2 && a.dpy == 2 && a.dpz == 2;
  const bool flat = a.d1y == 1 && a.d1z == 1;
  if (trilinear && flat &&
      (op == pp::OP_PULL || op == pp::OP_PULLG || op == pp::OP_PULLH || op == pp::OP_PULL_BWD ||
       op == pp::OP_PULLG_BWD))
    return launch_variant(op, a, nbatch, s, 0);
  return launch_variant(op, a, nbatch, s, tiled);
}

struct Geo {
  long long n0, n1;
};

// Validate a descriptor (src/pushpull_general.jl:26-57 normalises the same fields;
// here out-of-range values are rejected, not clamped).
int check_desc(const pp_desc* d, Geo& g) {
  if (!d) return fail(PP_EINVAL, "null descriptor");
  g.n0 = 1;
  g.n1 = 1;
  for (int a = 0; a < 3; ++a) {
    if (d->d0[a] < 1 || d->d1[a] < 1) return fail(PP_EINVAL, "grid sizes must be >= 1 (axis %d)", a);
    if (d->d0[a] > kMaxDim || d->d1[a] > kMaxDim) return fail(PP_ERANGE, "axis %d larger than 2^29", a);
    g.n0 *= d->d0[a];
    g.n1 *= d->d1[a];
    if (d->deg[a] < 0 || d->deg[a] > 4) return fail(PP_EINVAL, "degree %d on axis %d outside 0..4", d->deg[a], a);
  }
  if (g.n0 > kMaxVox || g.n1 > kMaxVox) return fail(PP_ERANGE, "a volume holds more than 2^31-1 voxels");
  if (d->nchan < 1) return fail(PP_EINVAL, "nchan must be >= 1");
  if (d->nbatch < 1 || d->nbatch > 65535) return fail(PP_EINVAL, "nbatch must be in 1..65535");
  if (!(d->nbnd == 1 || d->nbnd == d->nchan)) return fail(PP_EINVAL, "nbnd must be 1 or nchan");
  if (d->nbnd > PP_MAX_CHAN_BND) return fail(PP_EINVAL, "more than %d per-channel boundary triples", PP_MAX_CHAN_BND);
  for (int i = 0; i < 3 * d->nbnd; ++i)
    if (d->bnd[i] < 0 || d->bnd[i] > 2) return fail(PP_EINVAL, "boundary code %d not in {0,1,2}", d->bnd[i]);
  return PP_OK;
}

void base_args(const pp_desc* d, pp::KArgs& a, const Geo& g) {
  std::memset(&a, 0, sizeof a);
  a.n1 = (int)g.n1;
  a.d1x = (int)d->d1[0];
  a.d1y = (int)d->d1[1];
  a.d1z = (int)d->d1[2];
  a.d0x = (int)d->d0[0];
  a.d0y = (int)d->d0[1];
  a.d0z = (int)d->d0[2];
  a.dpx = d->deg[0] + 1;
  a.dpy = d->deg[1] + 1;
  a.dpz = d->deg[2] + 1;
  a.ext = d->ext != 0;
  a.lx = (float)d->d0[0] - 0.99f;
  a.ly = (float)d->d0[1] - 0.99f;
  a.lz = (float)d->d0[2] - 0.99f;
  a.phi_bs = 3 * g.n1;
  a.vscale = 1.0f;
}

bool uniform_bnd(const pp_desc* d) {
  for (int c = 1; c < d->nbnd; ++c)
    for (int q = 0; q < 3; ++q)
      if (d->bnd[3 * c + q] != d->bnd[q]) return false;
  return true;
}

// Channels whose boundary codes differ only between reflect (1) and antisymmetric (2)
// fold to the same indices and differ only in sign. Returns true when d's per-channel
// codes are like that on every axis (or equal), with the shared fold codes in `base` and
// the per-channel sign bits (3c + axis) in `mask`.
bool sign_split(const pp_desc* d, int base[3], unsigned& mask) {
  if (d->nchan > 10 || d->nbnd != d->nchan) return false;
  mask = 0u;
  for (int q = 0; q < 3; ++q) {
    bool any2 = false, same = true, refl = true;
    for (int c = 0; c < d->nchan; ++c) {
      const int code = d->bnd[3 * c + q];
      any2 |= code == 2;
      same &= code == d->bnd[q];
      refl &= code == 1 || code == 2;
      if (code == 2) mask |= 1u << (3 * c + q);
    }
    if (!same && !refl) return false;
    base[q] = any2 ? 2 : d->bnd[q];
  }
  return true;
}

// Launch `op` over all channels: one launch when the boundary codes are shared (or differ
// only in sign, on the tiled kernels of an isotropic degree 1-4), else one per channel
// (pointers offset by the channel strides; channel-reducing outputs accumulate after the
// first launch).
int run(pp::Op op, const pp_desc* d, const Geo& g, pp::KArgs a, hipStream_t s) {
  const bool reduce =
      (op == pp::OP_PULL_BWD || op == pp::OP_PULLG_BWD || op == pp::OP_OBJGRAD || op == pp::OP_PULL_VJP);
  if (uniform_bnd(d)) {
    a.bx = d->bnd[0];
    a.by = d->bnd[1];
    a.bz = d->bnd[2];
    a.nchan = d->nchan;
    hipError_t e = launch(op, a, d->nbatch, s);
    return e == hipSuccess ? ok() : hip_fail(e, "kernel launch");
  }
  int base[3];
  unsigned mask;
  const bool iso_tiled = a.dpx == a.dpy && a.dpx == a.dpz && a.dpx >= 2 && a.dpx <= 5;
  const int tiling = g_tiling.load(std::memory_order_relaxed);
  // per-sample kernels where they are forced: only the trilinear pull has a signed one
  const bool signed_ps = tiling == 0 && op == pp::OP_PULL && a.dpx == 2 && iso_tiled;
  if (((iso_tiled && tiling != 0) || signed_ps) && sign_split(d, base, mask)) {
    pp::KArgs b = a;
    b.bx = base[0];
    b.by = base[1];
    b.bz = base[2];
    b.csign_on = 1;
    b.csign = mask;
    b.nchan = d->nchan;
    // the tiled kernels and the per-sample trilinear pull take the signs; an op the tiled
    // kernels decline runs once per channel below
    hipError_t e = signed_ps
                       ? pp::launch_op(op, b, d->nbatch, s, kBlock)
                       : pp::launch_tiled(op, b, d->nbatch, s,
                                          pp::TileCfg{g_shape.load(std::memory_order_relaxed),
                                                      g_cap.load(std::memory_order_relaxed)});
    if (e != hipErrorNotSupported) return e == hipSuccess ? ok() : hip_fail(e, "kernel launch");
  }
  for (int c = 0; c < d->nchan; ++c) {
    pp::KArgs b = a;
    b.bx = d->bnd[3 * c];
    b.by = d->bnd[3 * c + 1];
    b.bz = d->bnd[3 * c + 2];
    b.nchan = 1;
    b.chan0 = c;
    b.f = a.f + c * a.f_cs;
    if (a.aux) b.aux = a.aux + c * a.aux_cs;
    if (reduce) {
      b.accum = c > 0 ? 1 : a.accum;
      if (a.out2) b.out2 = a.out2 + c * a.out2_cs;
    } else {
      b.out = a.out + c * a.out_cs;
    }
    hipError_t e = launch(op, b, d->nbatch, s);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
  }
  (void)g;
  return ok();
}

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int pull_family(pp::Op op, int ncomp, const pp_desc* d, float* out, const float* phi, const float* f0, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!out || !phi || !f0) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{out, "out"}, {phi, "phi"}, {f0, "f0"}})) return r;
  pp::KArgs a;
  base_args(d, a, g);
  a.f = f0;
  a.f_cs = g.n0;
  a.f_bs = g.n0 * d->nchan;
  a.phi = phi;
  a.out = out;
  a.out_cs = ncomp * g.n1;
  a.out_bs = ncomp * g.n1 * d->nchan;
  return run(op, d, g, a, S(s));
}

int push_family(pp::Op op, int ncomp, const pp_desc* d, float* f0, const float* phi, const float* f1, int accumulate,
                void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!f0 || !phi || !f1) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{f0, "f0"}, {phi, "phi"}, {f1, "f1"}})) return r;
  if (!accumulate) {
    hipError_t e = pp::launch_fill(f0, g.n0 * d->nchan * d->nbatch, 0.0f, S(s));
    if (e != hipSuccess) return hip_fail(e, "zero fill");
  }
  pp::KArgs a;
  base_args(d, a, g);
  a.f = f1;
  a.f_cs = ncomp * g.n1;
  a.f_bs = ncomp * g.n1 * d->nchan;
  a.phi = phi;
  a.out = f0;
  a.out_cs = g.n0;
  a.out_bs = g.n0 * d->nchan;
  return run(op, d, g, a, S(s));
}

int backward_family(pp::Op op, int ncomp, const pp_desc* d, float* dphi, const float* phi, const float* f0,
                    const float* delta, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!dphi || !phi || !f0 || !delta) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{dphi, "dphi"}, {phi, "phi"}, {f0, "f0"}, {delta, "delta"}})) return r;
  pp::KArgs a;
  base_args(d, a, g);
  a.f = f0;
  a.f_cs = g.n0;
  a.f_bs = g.n0 * d->nchan;
  a.phi = phi;
  a.aux = delta;
  a.aux_cs = ncomp * g.n1;
  a.aux_bs = ncomp * g.n1 * d->nchan;
  a.out = dphi;
  a.out_bs = 3 * g.n1;
  return run(op, d, g, a, S(s));
}

// ---- Tier A: host pointers, one slice, synchronous ---------------------------
struct Staging {
  std::mutex mu;
  void* buf = nullptr;
  size_t cap = 0;
  int dev = -1;
};
Staging g_stage[64];

// Checks a Tier A call and fills a one-slice descriptor view.
int tierA_args(const pp_usize* d0, pp_usize n1, const int* bnd, const pp_usize* dp, int ext, pp::KArgs& a,
               long long& n0) {
  if (!d0 || !bnd || !dp) return fail(PP_EINVAL, "null d0/bnd/dp");
  n0 = 1;
  for (int q = 0; q < 3; ++q) {
    if (d0[q] < 1 || d0[q] > (pp_usize)kMaxDim) return fail(PP_EINVAL, "d0[%d] out of range", q);
    if (dp[q] < 1 || dp[q] > 5) return fail(PP_EINVAL, "dp[%d]=%llu outside 1..5", q, (unsigned long long)dp[q]);
    n0 *= (long long)d0[q];
  }
  if (n0 > kMaxVox || n1 > (pp_usize)kMaxVox) return fail(PP_ERANGE, "slice larger than 2^31-1 voxels");
  std::memset(&a, 0, sizeof a);
  a.n1 = (int)n1;
  a.d1x = (int)n1;  // Tier A sees a flat list of samples
  a.d1y = 1;
  a.d1z = 1;
  a.d0x = (int)d0[0];
  a.d0y = (int)d0[1];
  a.d0z = (int)d0[2];
  a.dpx = (int)dp[0];
  a.dpy = (int)dp[1];
  a.dpz = (int)dp[2];
  a.bx = bnd[0];
  a.by = bnd[1];
  a.bz = bnd[2];
  a.ext = ext != 0;
  a.lx = (float)d0[0] - 0.99f;
  a.ly = (float)d0[1] - 0.99f;
  a.lz = (float)d0[2] - 0.99f;
  a.nchan = 1;
  a.phi_bs = 3 * (long long)n1;
  a.vscale = 1.0f;
  return PP_OK;
}

// in_n: floats of the gathered/pushed input; out_n: floats of the output; the output is
// uploaded first when `upload_out` (push accumulates into the caller's array).
// aff/d1: affine variant (pushpull.c:176-244): coordinates from the 0-based 3x4 matrix
// over the output grid d1 instead of from phi.
int tierA(pp::Op op, float* out, long long out_n, bool upload_out, const float* phi, const float* in, long long in_n,
          const pp_usize* d0, pp_usize n1, const int* bnd, const pp_usize* dp, int ext,
          const float* aff = nullptr, const pp_usize* d1 = nullptr) {
  pp::KArgs a;
  long long n0;
  if (int r = tierA_args(d0, n1, bnd, dp, ext, a, n0)) return r;
  if (aff) {
    if (!d1) return fail(PP_EINVAL, "null d1");
    long long m = 1;
    for (int q = 0; q < 3; ++q) {
      if (d1[q] < 1 || d1[q] > (pp_usize)kMaxDim) return fail(PP_EINVAL, "d1[%d] out of range", q);
      m *= (long long)d1[q];
    }
    if (m > kMaxVox) return fail(PP_ERANGE, "output grid larger than 2^31-1 voxels");
    // the reference loops over d1 (pushpull.c:185-208); n1 only sizes the caller's buffer
    n1 = (pp_usize)m;
    a.n1 = (int)m;
    a.d1x = (int)d1[0];
    a.d1y = (int)d1[1];
    a.d1z = (int)d1[2];
    a.affine = 1;
    for (int q = 0; q < 12; ++q) a.aff[q] = aff[q];
    phi = in;  // no deformation array is uploaded
  }
  if (n1 == 0) return ok();
  if (!out || !phi || !in) return fail(PP_EINVAL, "null array");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  Staging& st = g_stage[dev & 63];
  std::lock_guard<std::mutex> lock(st.mu);
  const size_t phi_b = aff ? 0 : 3 * (size_t)n1 * 4, in_b = (size_t)in_n * 4, out_b = (size_t)out_n * 4;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t need = al(phi_b) + al(in_b) + al(out_b);
  if (st.cap < need || st.dev != dev) {
    if (st.buf) (void)hipFree(st.buf);
    st.buf = nullptr;
    st.cap = 0;
    e = hipMalloc(&st.buf, need);
    if (e != hipSuccess) return fail(PP_ENOMEM, "hipMalloc(%zu): %s", need, hipGetErrorString(e));
    st.cap = need;
    st.dev = dev;
  }
  char* base = static_cast<char*>(st.buf);
  float* dphi = reinterpret_cast<float*>(base);
  float* din = reinterpret_cast<float*>(base + al(phi_b));
  float* dout = reinterpret_cast<float*>(base + al(phi_b) + al(in_b));
  hipStream_t s = nullptr;
  if (phi_b && (e = hipMemcpyAsync(dphi, phi, phi_b, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "upload phi");
  if ((e = hipMemcpyAsync(din, in, in_b, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail(e, "upload input");
  if (upload_out) {
    if ((e = hipMemcpyAsync(dout, out, out_b, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail(e, "upload output");
  }
  a.phi = dphi;
  a.f = din;
  a.out = dout;
  // one channel, one batch: channel/batch strides are never applied
  a.f_cs = in_n;
  a.out_cs = out_n;
  e = aff ? launch(op, a, 1, s) : pp::launch_op(op, a, 1, s, kBlock);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  if ((e = hipMemcpyAsync(out, dout, out_b, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail(e, "download");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "synchronize");
  return ok();
}

void nan_fill(float* p, long long n) {
  if (!p) return;
  for (long long i = 0; i < n; ++i) p[i] = NAN;
}

long long prod3(const pp_usize* d) { return d ? (long long)(d[0] * d[1] * d[2]) : 0; }

}  // namespace

namespace pp {
int abi_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int abi_ok() { return ok(); }
}  // namespace pp

extern "C" {

// ---- Tier A --------------------------------------------------------------------
void pull(float* f1, const float* phi, const float* f0, const pp_usize* d0, pp_usize n1, const int* bnd,
          const pp_usize* dp, int ext) {
  if (tierA(pp::OP_PULL, f1, (long long)n1, false, phi, f0, prod3(d0), d0, n1, bnd, dp, ext))
    nan_fill(f1, (long long)n1);
}

void pullg(float* g1, const float* phi, const float* f0, const pp_usize* d0, pp_usize n1, const int* bnd,
           const pp_usize* dp, int ext) {
  if (tierA(pp::OP_PULLG, g1, 3 * (long long)n1, false, phi, f0, prod3(d0), d0, n1, bnd, dp, ext))
    nan_fill(g1, 3 * (long long)n1);
}

void pullh(float* h1, const float* phi, const float* f0, const pp_usize* d0, pp_usize n1, const int* bnd,
           const pp_usize* dp, int ext) {
  if (tierA(pp::OP_PULLH, h1, 9 * (long long)n1, false, phi, f0, prod3(d0), d0, n1, bnd, dp, ext))
    nan_fill(h1, 9 * (long long)n1);
}

void push(float* f0, const float* phi, const float* f1, const pp_usize* d0, pp_usize n1, const int* bnd,
          const pp_usize* dp, int ext) {
  (void)tierA(pp::OP_PUSH, f0, prod3(d0), true, phi, f1, (long long)n1, d0, n1, bnd, dp, ext);
}

void pushg(float* f0, const float* phi, const float* g1, const pp_usize* d0, pp_usize n1, const int* bnd,
           const pp_usize* dp, int ext) {
  (void)tierA(pp::OP_PUSHG, f0, prod3(d0), true, phi, g1, 3 * (long long)n1, d0, n1, bnd, dp, ext);
}

void pull_affine(float* f1, const float* f0, const pp_usize* d0, pp_usize n1, const pp_usize* d1, const float* Aff,
                 const int* bnd, const pp_usize* dp, int ext) {
  const long long m = prod3(d1);
  if (tierA(pp::OP_PULL, f1, m, false, nullptr, f0, prod3(d0), d0, n1, bnd, dp, ext, Aff, d1)) nan_fill(f1, m);
}

void push_affine(float* f0, const float* f1, const pp_usize* d0, pp_usize n1, const pp_usize* d1, const float* Aff,
                 const int* bnd, const pp_usize* dp, int ext) {
  (void)tierA(pp::OP_PUSH, f0, prod3(d0), true, nullptr, f1, prod3(d1), d0, n1, bnd, dp, ext, Aff, d1);
}

// ---- Tier B --------------------------------------------------------------------
int pp_pull(const pp_desc* d, float* f1, const float* phi, const float* f0, void* s) {
  return pull_family(pp::OP_PULL, 1, d, f1, phi, f0, s);
}
int pp_pull_grad(const pp_desc* d, float* g1, const float* phi, const float* f0, void* s) {
  return pull_family(pp::OP_PULLG, 3, d, g1, phi, f0, s);
}
int pp_pull_hess(const pp_desc* d, float* h1, const float* phi, const float* f0, void* s) {
  return pull_family(pp::OP_PULLH, 9, d, h1, phi, f0, s);
}
int pp_push(const pp_desc* d, float* f0, const float* phi, const float* f1, int accumulate, void* s) {
  return push_family(pp::OP_PUSH, 1, d, f0, phi, f1, accumulate, s);
}
int pp_push_grad(const pp_desc* d, float* f0, const float* phi, const float* g1, int accumulate, void* s) {
  return push_family(pp::OP_PUSHG, 3, d, f0, phi, g1, accumulate, s);
}
int pp_affine_pull(const pp_desc* d, float* f1, const float* f0, const float* aff, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!f1 || !f0 || !aff) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{f1, "f1"}, {f0, "f0"}})) return r;
  pp::KArgs a;
  base_args(d, a, g);
  a.affine = 1;
  for (int q = 0; q < 12; ++q) a.aff[q] = aff[q];
  a.phi = f0;  // unused: coordinates come from the affine
  a.phi_bs = 0;
  a.f = f0;
  a.f_cs = g.n0;
  a.f_bs = g.n0 * d->nchan;
  a.out = f1;
  a.out_cs = g.n1;
  a.out_bs = g.n1 * d->nchan;
  return run(pp::OP_PULL, d, g, a, S(s));
}

int pp_affine_push(const pp_desc* d, float* f0, const float* f1, const float* aff, int accumulate, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!f0 || !f1 || !aff) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{f0, "f0"}, {f1, "f1"}})) return r;
  if (!accumulate) {
    hipError_t e = pp::launch_fill(f0, g.n0 * d->nchan * d->nbatch, 0.0f, S(s));
    if (e != hipSuccess) return hip_fail(e, "zero fill");
  }
  pp::KArgs a;
  base_args(d, a, g);
  a.affine = 1;
  for (int q = 0; q < 12; ++q) a.aff[q] = aff[q];
  a.phi = f1;
  a.phi_bs = 0;
  a.f = f1;
  a.f_cs = g.n1;
  a.f_bs = g.n1 * d->nchan;
  a.out = f0;
  a.out_cs = g.n0;
  a.out_bs = g.n0 * d->nchan;
  return run(pp::OP_PUSH, d, g, a, S(s));
}

int pp_pull_backward(const pp_desc* d, float* dphi, const float* phi, const float* f0, const float* delta, void* s) {
  return backward_family(pp::OP_PULL_BWD, 1, d, dphi, phi, f0, delta, s);
}
int pp_pull_grad_backward(const pp_desc* d, float* dphi, const float* phi, const float* f0, const float* delta,
                          void* s) {
  return backward_family(pp::OP_PULLG_BWD, 3, d, dphi, phi, f0, delta, s);
}

int pp_objective_grad(const pp_desc* d, float* dphi, float* df0, double* loss, const float* phi, const float* f0,
                      const float* gt, int accumulate, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!dphi || !df0 || !loss || !phi || !f0 || !gt) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{dphi, "dphi"}, {df0, "df0"}, {loss, "loss"}, {phi, "phi"}, {f0, "f0"}, {gt, "g"}})) return r;
  if (!accumulate) {
    hipError_t e = pp::launch_fill(df0, g.n0 * d->nchan * d->nbatch, 0.0f, S(s));
    if (e == hipSuccess) e = pp::launch_fill_d(loss, 1, 0.0, S(s));
    if (e != hipSuccess) return hip_fail(e, "zero fill");
  }
  pp::KArgs a;
  base_args(d, a, g);
  a.f = f0;
  a.f_cs = g.n0;
  a.f_bs = g.n0 * d->nchan;
  a.phi = phi;
  a.aux = gt;
  a.aux_cs = g.n1;
  a.aux_bs = g.n1 * d->nchan;
  a.out = dphi;
  a.out_bs = 3 * g.n1;
  a.out2 = df0;
  a.out2_cs = g.n0;
  a.out2_bs = g.n0 * d->nchan;
  a.red = loss;
  return run(pp::OP_OBJGRAD, d, g, a, S(s));
}

int pp_objective_grad_step(const pp_desc* d, float* dphi, float* df0, double* loss, const float* phi,
                           const float* f0, const float* gt, float* phi_next, float step, int accumulate, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!phi_next) return fail(PP_EINVAL, "null phi_next");
  if (int r = check_mem({{phi_next, "phi_next"}})) return r;
  const long long n = 3 * g.n1 * d->nbatch;
  auto ovl = [n](const float* p, const float* q) { return q && p < q + n && q < p + n; };
  if (phi_next != phi && ovl(phi_next, phi)) return fail(PP_EINVAL, "phi_next overlaps phi without being it");
  if (ovl(phi_next, dphi)) return fail(PP_EINVAL, "phi_next must not overlap dphi");
  // One tiled launch covering every channel takes the step once a tile's dphi is final:
  // shared boundary codes, or codes that differ only in sign.
  int base[3];
  unsigned mask = 0u;
  const bool uni_b = uniform_bnd(d);
  const bool tiled_ok = g_tiling.load(std::memory_order_relaxed) != 0 &&
                        (uni_b || (d->deg[0] == d->deg[1] && d->deg[0] == d->deg[2] && sign_split(d, base, mask)));
  if (tiled_ok) {
    if (!dphi || !df0 || !loss || !phi || !f0 || !gt) return fail(PP_EINVAL, "null array");
    if (int r = check_mem({{dphi, "dphi"}, {df0, "df0"}, {loss, "loss"}, {phi, "phi"}, {f0, "f0"}, {gt, "g"}}))
      return r;
    pp::KArgs a;
    base_args(d, a, g);
    a.f = f0;
    a.f_cs = g.n0;
    a.f_bs = g.n0 * d->nchan;
    a.phi = phi;
    a.aux = gt;
    a.aux_cs = g.n1;
    a.aux_bs = g.n1 * d->nchan;
    a.out = dphi;
    a.out_bs = 3 * g.n1;
    a.out2 = df0;
    a.out2_cs = g.n0;
    a.out2_bs = g.n0 * d->nchan;
    a.red = loss;
    a.nchan = d->nchan;
    a.bx = uni_b ? d->bnd[0] : base[0];
    a.by = uni_b ? d->bnd[1] : base[1];
    a.bz = uni_b ? d->bnd[2] : base[2];
    a.csign_on = uni_b ? 0 : 1;
    a.csign = uni_b ? 0u : mask;
    a.phi_next = phi_next;
    a.step = step;
    hipError_t e = hipSuccess;
    if (!accumulate) {
      e = pp::launch_fill(df0, g.n0 * d->nchan * d->nbatch, 0.0f, S(s));
      if (e == hipSuccess) e = pp::launch_fill_d(loss, 1, 0.0, S(s));
    }
    if (e == hipSuccess)
      e = pp::launch_tiled(pp::OP_OBJGRAD, a, d->nbatch, S(s),
                           pp::TileCfg{g_shape.load(std::memory_order_relaxed), g_cap.load(std::memory_order_relaxed)});
    if (e != hipErrorNotSupported) return e == hipSuccess ? ok() : hip_fail(e, "kernel launch");
    // (degree 0 on every axis: no tiled kernel; the zero fills above are repeated below)
  }
  if (int r = pp_objective_grad(d, dphi, df0, loss, phi, f0, gt, accumulate, s)) return r;
  hipError_t e = hipSuccess;
  if (phi_next != phi) e = hipMemcpyAsync(phi_next, phi, sizeof(float) * n, hipMemcpyDeviceToDevice, S(s));
  if (e == hipSuccess) e = pp::launch_axpy(phi_next, dphi, step, n, S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "descent step");
}

int pp_pull_vjp(const pp_desc* d, float* dphi, float* df0, const float* phi, const float* f0, const float* delta,
                int accumulate, void* s) {
  Geo g;
  if (int r = check_desc(d, g)) return r;
  if (!dphi || !df0 || !phi || !f0 || !delta) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{dphi, "dphi"}, {df0, "df0"}, {phi, "phi"}, {f0, "f0"}, {delta, "delta"}})) return r;
  if (!accumulate) {
    hipError_t e = pp::launch_fill(df0, g.n0 * d->nchan * d->nbatch, 0.0f, S(s));
    if (e != hipSuccess) return hip_fail(e, "zero fill");
  }
  pp::KArgs a;
  base_args(d, a, g);
  a.f = f0;
  a.f_cs = g.n0;
  a.f_bs = g.n0 * d->nchan;
  a.phi = phi;
  a.aux = delta;
  a.aux_cs = g.n1;
  a.aux_bs = g.n1 * d->nchan;
  a.out = dphi;
  a.out_bs = 3 * g.n1;
  a.out2 = df0;
  a.out2_cs = g.n0;
  a.out2_bs = g.n0 * d->nchan;
  return run(pp::OP_PULL_VJP, d, g, a, S(s));
}

// ---- push with a count image ----------------------------------------------------------
namespace {

int push_count_impl(const pp_desc* d, float* f0, float* c0, const float* phi, const float* f1, const float* aff,
                    int accumulate, void* s) {
  if (!d) return fail(PP_EINVAL, "null descriptor");
  pp_desc dc = *d;  // the channels' triples only
  int cnt[3];
  if (d->nchan >= 1 && d->nbnd == d->nchan + 1 && d->nbnd <= PP_MAX_CHAN_BND) {
    for (int q = 0; q < 3; ++q) cnt[q] = d->bnd[3 * d->nchan + q];
    dc.nbnd = d->nchan;
  } else {
    for (int q = 0; q < 3; ++q) cnt[q] = d->bnd[q];
  }
  for (int q = 0; q < 3; ++q)
    if (cnt[q] < 0 || cnt[q] > 2) return fail(PP_EINVAL, "count boundary code %d not in {0,1,2}", cnt[q]);
  Geo g;
  if (int r = check_desc(&dc, g)) return r;
  if (!f0 || !c0 || (!aff && !phi)) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{f0, "f0"}, {c0, "c0"}, {phi, "phi"}, {f1, "f1"}})) return r;
  if (!f1 && d->nchan != 3) return fail(PP_EINVAL, "pushing the identity needs nchan == 3, not %d", d->nchan);
  const int nc = d->nchan, nb = d->nbatch;
  if (!accumulate) {
    hipError_t e = pp::launch_fill(f0, g.n0 * nc * nb, 0.0f, S(s));
    if (e == hipSuccess) e = pp::launch_fill(c0, g.n0 * nb, 0.0f, S(s));
    if (e != hipSuccess) return hip_fail(e, "zero fill");
  }
  pp::KArgs a;
  base_args(&dc, a, g);
  a.f = f1;
  a.f_cs = g.n1;
  a.f_bs = g.n1 * nc;
  a.vmode = f1 ? 0 : 2;
  if (aff) {
    a.affine = 1;
    for (int q = 0; q < 12; ++q) a.aff[q] = aff[q];
    a.phi = c0;  // not read: coordinates come from the matrix
    a.phi_bs = 0;
  } else {
    a.phi = phi;
  }
  a.out = f0;
  a.out_cs = g.n0;
  a.out_bs = g.n0 * nc;
  // one launch: channels and count fold alike and the tiled push covers the degree
  pp_desc dd = dc;
  dd.nchan = nc + 1;
  dd.nbnd = nc + 1;
  const bool fits = nc + 1 <= PP_MAX_CHAN_BND;
  if (fits) {
    for (int c = 0; c < nc; ++c)
      for (int q = 0; q < 3; ++q) dd.bnd[3 * c + q] = dc.bnd[3 * (dc.nbnd == 1 ? 0 : c) + q];
    for (int q = 0; q < 3; ++q) dd.bnd[3 * nc + q] = cnt[q];
  }
  int base[3];
  unsigned mask;
  const bool iso_tiled = a.dpx == a.dpy && a.dpx == a.dpz && a.dpx >= 2 && a.dpx <= 5;
  if (fits && iso_tiled && g_tiling.load(std::memory_order_relaxed) != 0 && sign_split(&dd, base, mask)) {
    a.bx = base[0];
    a.by = base[1];
    a.bz = base[2];
    a.csign_on = 1;
    a.csign = mask;
    a.nchan = nc;
    a.out2 = c0;
    a.out2_bs = g.n0;
    hipError_t e = launch_variant(pp::OP_PUSH, a, nb, S(s), 1 + g_shape.load(std::memory_order_relaxed));
    return e == hipSuccess ? ok() : hip_fail(e, "kernel launch");
  }
  if (int r = run(pp::OP_PUSH, &dc, g, a, S(s))) return r;
  pp::KArgs b = a;  // the count: ones through the count's triple
  b.vmode = 1;
  b.nchan = 1;
  b.chan0 = 0;
  b.bx = cnt[0];
  b.by = cnt[1];
  b.bz = cnt[2];
  b.out = c0;
  b.out_cs = g.n0;
  b.out_bs = g.n0;
  hipError_t e = launch(pp::OP_PUSH, b, nb, S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "kernel launch");
}

}  // namespace

int pp_push_count(const pp_desc* d, float* f0, float* c0, const float* phi, const float* f1, int accumulate,
                  void* s) {
  return push_count_impl(d, f0, c0, phi, f1, nullptr, accumulate, s);
}

int pp_affine_push_count(const pp_desc* d, float* f0, float* c0, const float* f1, const float* aff, int accumulate,
                         void* s) {
  if (!aff) return fail(PP_EINVAL, "null affine");
  return push_count_impl(d, f0, c0, nullptr, f1, aff, accumulate, s);
}

// ---- compositions (src/compositions.jl) ---------------------------------------------
namespace {

// Descriptor of compose: 3 channels, degree 1, extrapolation, boundary (1,1,1) with 2 on
// the channel's own axis (compositions.jl:32-35).
void compose_desc(pp_desc& d, const long long d0[3], const long long d1[3], int nbatch) {
  std::memset(&d, 0, sizeof d);
  for (int q = 0; q < 3; ++q) {
    d.d0[q] = d0[q];
    d.d1[q] = d1[q];
    d.deg[q] = 1;
  }
  d.nchan = 3;
  d.nbatch = nbatch;
  d.ext = 1;
  d.nbnd = 3;
  for (int c = 0; c < 3; ++c)
    for (int q = 0; q < 3; ++q) d.bnd[3 * c + q] = (q == c) ? 2 : 1;
}

bool overlap(const float* p, const float* q, long long n) { return p < q + n && q < p + n; }

}  // namespace

int pp_compose(float* psi, const float* phi, const float* theta, const long long d0[3], const long long d1[3],
               int nbatch, void* s) {
  if (!psi || !phi || !theta || !d0 || !d1) return fail(PP_EINVAL, "null argument");
  if (int r = check_mem({{psi, "psi"}, {phi, "phi"}, {theta, "theta"}})) return r;
  pp_desc d;
  compose_desc(d, d0, d1, nbatch);
  Geo g;
  if (int r = check_desc(&d, g)) return r;
  if (overlap(psi, phi, 3 * g.n0 * nbatch) || overlap(psi, theta, 3 * g.n1 * nbatch))
    return fail(PP_EINVAL, "psi must not overlap phi or theta");
  return pull_family(pp::OP_PULL, 1, &d, psi, theta, phi, s);
}

int pp_exp(float* out, const float* v, float* work, const long long dim[3], int nbatch, int nsq, void* s) {
  if (!out || !v || !work || !dim) return fail(PP_EINVAL, "null argument");
  if (int r = check_mem({{out, "out"}, {v, "v"}, {work, "work"}})) return r;
  if (nsq < 1 || nsq > 30) return fail(PP_EINVAL, "nsq must be in 1..30, not %d", nsq);
  pp_desc d;
  compose_desc(d, dim, dim, nbatch);
  Geo g;
  if (int r = check_desc(&d, g)) return r;
  const long long n = 3 * g.n1 * nbatch;
  if (overlap(out, v, n) || overlap(work, v, n) || overlap(out, work, n))
    return fail(PP_EINVAL, "out, v and work must not overlap");
  pp::KArgs a;
  base_args(&d, a, g);
  a.disp = 1;
  a.f_cs = g.n1;
  a.f_bs = 3 * g.n1;
  a.out_cs = g.n1;
  a.out_bs = 3 * g.n1;
  const float* src = v;
  for (int k = 1; k <= nsq; ++k) {
    // the last squaring lands in `out`: alternate so that (nsq - k) even writes out
    float* dst = ((nsq - k) % 2 == 0) ? out : work;
    a.f = src;
    a.phi = src;
    a.out = dst;
    // first squaring: u = v / 2^nsq, applied to the displacement and the gathered values
    a.dscale = a.vscale = (k == 1) ? std::ldexp(1.0f, -nsq) : 1.0f;
    a.add_id = k == nsq;  // u .+= Id
    if (int r = run(pp::OP_PULL, &d, g, a, S(s))) return r;
    src = dst;
  }
  return ok();
}

// ---- batches of uneven volumes -----------------------------------------------------
struct pp_batch_group {
  int dp;      // degree + 1 of every volume in the group (the largest of its axes)
  int an;      // 1: the group's degrees differ between axes
  int first;   // index of the group's first volume in the device array
  int nvol;
  int total;   // tiles over the group's volumes
  pp::XcdSplit split;  // per-XCD tile ranges of equal work
};

struct pp_batch_plan {
  pp::Op op;
  int nvol;
  std::vector<pp_batch_group> groups;  // volumes ordered by degree group, one launch per group
  pp::BatchVol* dvols;
  // Degree groups after the first run beside it: one stream each (default priority),
  // forked from and joined back to the caller's stream with events, so that the groups'
  // workgroups share the device and a short group fills the tail of a long one. Made
  // once with the plan; graph capture records the fork and join as graph edges.
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
};

namespace {

// A volume's degree group: its largest degree, and whether its degrees differ between axes
// (those volumes run on kernels of their own). Groups 0-4: one degree on every axis; 6-9:
// degrees that differ, the largest 1-4.
int deg_group(const pp_volume& v) {
  const int lo = std::min(v.deg[0], std::min(v.deg[1], v.deg[2]));
  const int hi = std::max(v.deg[0], std::max(v.deg[1], v.deg[2]));
  return hi + (lo != hi ? 5 : 0);
}
constexpr int kDegGroups = 10;

// Volume i's boundary codes into a: one triple for every channel (nbnd 0 or 1), or a triple
// per channel (nbnd = nchan) whose codes along each axis are equal or all reflect (1) /
// antisymmetric (2): those fold to the same indices and differ in sign only, so all
// channels fold with one code per axis and each takes its own signs (KArgs::csign).
int batch_bnd(const pp_volume& v, int i, pp::KArgs& a) {
  const int nb = v.nbnd;
  if (nb < 0 || (nb > 1 && nb != v.nchan))
    return fail(PP_EINVAL, "volume %d: nbnd must be 0, 1 or nchan (%d), got %d", i, v.nchan, nb);
  if (nb > PP_BATCH_CHAN_BND)
    return fail(PP_EINVAL, "volume %d: at most %d channels take boundary codes of their own, got %d", i,
                PP_BATCH_CHAN_BND, nb);
  const int nc = nb > 1 ? nb : 1;
  const int32_t* codes = nb > 1 ? v.chan_bnd : v.bnd;
  int base[3];
  unsigned mask = 0u;
  bool split = false;
  for (int q = 0; q < 3; ++q) {
    bool any2 = false, same = true, refl = true;
    for (int c = 0; c < nc; ++c) {
      const int code = codes[3 * c + q];
      if (code < 0 || code > 2) return fail(PP_EINVAL, "volume %d: boundary code %d", i, code);
      any2 |= code == 2;
      same &= code == codes[q];
      refl &= code == 1 || code == 2;
      if (code == 2) mask |= 1u << (3 * c + q);
    }
    if (!same && !refl)
      return fail(PP_EINVAL,
                  "volume %d: the channels' boundary codes along axis %d differ other than between 1 and 2 "
                  "(run such a volume with single calls)", i, q);
    split |= !same;
    base[q] = any2 ? 2 : codes[q];
  }
  a.bx = base[0];
  a.by = base[1];
  a.bz = base[2];
  a.csign_on = split ? 1 : 0;
  a.csign = split ? mask : 0u;
  return PP_OK;
}

// Validate a batch and lay it out: hv[0..nvol) receives the volumes' kernel arguments
// ordered by degree group (stable: input order within a group), groups[0..*ngroups) one
// launch each (at most nine groups). Writes only into the caller's arrays (no allocation).
int build_batch(int op, const pp_volume* vols, int nvol, pp::BatchVol* hv, pp_batch_group* groups, int* ngroups) {
  if (!vols || nvol < 1) return fail(PP_EINVAL, "need at least one volume");
  if (op != PP_OP_PULL && op != PP_OP_PUSH && op != PP_OP_PUSH_GRAD && op != PP_OP_PULL_BACKWARD &&
      op != PP_OP_OBJECTIVE && op != PP_OP_PULL_VJP)
    return fail(PP_EINVAL, "unsupported batch op %d", op);
  for (int i = 0; i < nvol; ++i)
    for (int q = 0; q < 3; ++q)
      if (vols[i].deg[q] < 0 || vols[i].deg[q] > 4)
        return fail(PP_EINVAL, "volume %d: degree %d outside 0..4 on axis %d", i, vols[i].deg[q], q);
  long long total = 0;
  int pos = 0;
  for (int grp = 0; grp < kDegGroups; ++grp)
  for (int i = 0; i < nvol; ++i) {
    const pp_volume& v = vols[i];
    if (deg_group(v) != grp) continue;
    long long n0 = 1, n1 = 1;
    for (int q = 0; q < 3; ++q) {
      if (v.d0[q] < 1 || v.d1[q] < 1 || v.d0[q] > kMaxDim || v.d1[q] > kMaxDim)
        return fail(PP_EINVAL, "volume %d: grid size out of range on axis %d", i, q);
      n0 *= v.d0[q];
      n1 *= v.d1[q];
    }
    if (n0 > kMaxVox || n1 > kMaxVox) return fail(PP_ERANGE, "volume %d holds more than 2^31-1 voxels", i);
    if (v.nchan < 1) return fail(PP_EINVAL, "volume %d: nchan must be >= 1", i);
    pp::KArgs codes{};
    if (int r = batch_bnd(v, i, codes)) return r;
    if (!v.f || !v.phi || !v.out) return fail(PP_EINVAL, "volume %d: null array", i);
    if ((op == PP_OP_PULL_BACKWARD || op == PP_OP_OBJECTIVE || op == PP_OP_PULL_VJP) && !v.aux)
      return fail(PP_EINVAL, "volume %d: null aux", i);
    if (op == PP_OP_OBJECTIVE && (!v.out2 || !v.loss)) return fail(PP_EINVAL, "volume %d: null out2/loss", i);
    if (op == PP_OP_PULL_VJP && !v.out2) return fail(PP_EINVAL, "volume %d: null out2", i);
    if (v.phi_next && op != PP_OP_OBJECTIVE)
      return fail(PP_EINVAL, "volume %d: phi_next (a fused descent step) is only for PP_OP_OBJECTIVE", i);
    for (const Arr& m : {Arr{v.f, "f"}, Arr{v.phi, "phi"}, Arr{v.aux, "aux"}, Arr{v.out, "out"}, Arr{v.out2, "out2"},
                         Arr{v.loss, "loss"}, Arr{v.phi_next, "phi_next"}}) {
      if (int r = check_mem({m})) {
        char why[sizeof t_err];
        std::snprintf(why, sizeof why, "%s", t_err);
        return fail(r, "volume %d: %s", i, why);
      }
    }
    pp::BatchVol& bv = hv[pos++];
    pp::KArgs& a = bv.a;
    std::memset(&a, 0, sizeof a);
    a.n1 = (int)n1;
    a.d1x = (int)v.d1[0];
    a.d1y = (int)v.d1[1];
    a.d1z = (int)v.d1[2];
    a.d0x = (int)v.d0[0];
    a.d0y = (int)v.d0[1];
    a.d0z = (int)v.d0[2];
    a.dpx = v.deg[0] + 1;
    a.dpy = v.deg[1] + 1;
    a.dpz = v.deg[2] + 1;
    a.bx = codes.bx;
    a.by = codes.by;
    a.bz = codes.bz;
    a.csign_on = codes.csign_on;
    a.csign = codes.csign;
    a.ext = v.ext != 0;
    a.lx = (float)v.d0[0] - 0.99f;
    a.ly = (float)v.d0[1] - 0.99f;
    a.lz = (float)v.d0[2] - 0.99f;
    a.nchan = v.nchan;
    a.vscale = 1.0f;
    a.phi = v.phi;
    a.phi_bs = 3 * n1;
    a.f = v.f;
    a.out = v.out;
    bv.zero_n = 0;
    switch (op) {
      case PP_OP_PULL: a.f_cs = n0; a.out_cs = n1; break;
      case PP_OP_PUSH: a.f_cs = n1; a.out_cs = n0; bv.zero_n = n0 * v.nchan; break;
      case PP_OP_PUSH_GRAD: a.f_cs = 3 * n1; a.out_cs = n0; bv.zero_n = n0 * v.nchan; break;
      case PP_OP_PULL_BACKWARD: a.f_cs = n0; a.aux = v.aux; a.aux_cs = n1; break;
      default:
        a.f_cs = n0;
        a.aux = v.aux;
        a.aux_cs = n1;
        a.out2 = v.out2;
        a.out2_cs = n0;
        a.red = op == PP_OP_OBJECTIVE ? v.loss : nullptr;
        a.phi_next = op == PP_OP_OBJECTIVE ? v.phi_next : nullptr;
        a.step = v.step;
        bv.zero_n = n0 * v.nchan;
        break;
    }
    bv.dp = std::max(a.dpx, std::max(a.dpy, a.dpz));
    bv.an = grp >= 5;
    total += pp::batch_tiles(a);
    if (total > kMaxVox) return fail(PP_ERANGE, "too many tiles in one batch");
  }
  // number tiles within each degree group
  int ng = 0;
  for (int i = 0; i < nvol; ++i) {
    if (ng == 0 || groups[ng - 1].dp != hv[i].dp || groups[ng - 1].an != hv[i].an)
      groups[ng++] = pp_batch_group{hv[i].dp, hv[i].an, i, 0, 0, {}};
    pp_batch_group& g = groups[ng - 1];
    hv[i].tile0 = g.total;
    g.total += pp::batch_tiles(hv[i].a);
    g.nvol += 1;
  }
  // per-XCD ranges of equal work: a tile costs its volume's channel count plus one (the
  // per-tile setup); the cuts fall on tile boundaries
  for (int gi = 0; gi < ng; ++gi) {
    pp_batch_group& g = groups[gi];
    double work = 0.0;
    for (int i = g.first; i < g.first + g.nvol; ++i) work += (double)pp::batch_tiles(hv[i].a) * (hv[i].a.nchan + 1);
    int x = 1;
    double acc = 0.0;
    g.split.s[0] = 0;
    for (int i = g.first; i < g.first + g.nvol && x < pp::kBatchXcds; ++i) {
      const int nt = pp::batch_tiles(hv[i].a);
      const double c = hv[i].a.nchan + 1;
      while (x < pp::kBatchXcds && acc + nt * c >= work * x / pp::kBatchXcds) {
        const long long k = (long long)std::ceil((work * x / pp::kBatchXcds - acc) / c);
        g.split.s[x++] = hv[i].tile0 + (int)std::min<long long>(std::max<long long>(k, 0), nt);
      }
      acc += nt * c;
    }
    while (x <= pp::kBatchXcds) g.split.s[x++] = g.total;
    for (int q = 1; q <= pp::kBatchXcds; ++q) g.split.s[q] = std::max(g.split.s[q], g.split.s[q - 1]);
  }
  *ngroups = ng;
  return PP_OK;
}

// Zero fill and one persistent launch per degree group of a laid-out batch in device memory.
hipError_t run_batch(pp::Op op, const pp::BatchVol* dvols, int nvol, const pp_batch_group* groups, int ngroups,
                     hipStream_t s) {
  hipError_t e = pp::launch_batch_zero(op, dvols, nvol, s);
  for (int gi = 0; gi < ngroups && e == hipSuccess; ++gi)
    e = pp::launch_batch_group(op, groups[gi].dp, groups[gi].an != 0, dvols + groups[gi].first, groups[gi].nvol,
                               groups[gi].split, s);
  return e;
}

}  // namespace

int pp_batch_plan_create(int op, const pp_volume* vols, int nvol, pp_batch_plan** plan) {
  if (!plan) return fail(PP_EINVAL, "null plan pointer");
  *plan = nullptr;
  if (!vols || nvol < 1) return fail(PP_EINVAL, "need at least one volume");
  std::vector<pp::BatchVol> hv((size_t)nvol);
  pp_batch_group g[kDegGroups];
  int ng = 0;
  const int st = build_batch(op, vols, nvol, hv.data(), g, &ng);
  if (st != PP_OK) return st;
  pp::BatchVol* d = nullptr;
  hipError_t e = hipMalloc((void**)&d, sizeof(pp::BatchVol) * (size_t)nvol);
  if (e != hipSuccess) return fail(PP_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  e = hipMemcpy(d, hv.data(), sizeof(pp::BatchVol) * (size_t)nvol, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "upload batch descriptors");
  }
  pp_batch_plan* p = new pp_batch_plan{static_cast<pp::Op>(op), nvol, std::vector<pp_batch_group>(g, g + ng), d};
  // side streams and events for the degree groups after the first (default priority)
  bool made = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) == hipSuccess;
  for (int gi = 1; gi < ng && made; ++gi) {
    hipStream_t ss = nullptr;
    hipEvent_t ev = nullptr;
    made = hipStreamCreateWithFlags(&ss, hipStreamNonBlocking) == hipSuccess;
    if (made) p->side.push_back(ss);
    made = made && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    if (made) p->join.push_back(ev);
  }
  if (!made) {
    pp_batch_plan_destroy(p);
    return fail(PP_EDEVICE, "could not create the batch plan's side streams");
  }
  *plan = p;
  return ok();
}

int pp_batch_run(const pp_batch_plan* plan, void* s) {
  if (!plan) return fail(PP_EINVAL, "null plan");
  const int ng = (int)plan->groups.size();
  if (ng <= 1 || plan->side.size() + 1 < (size_t)ng) {
    const hipError_t e = run_batch(plan->op, plan->dvols, plan->nvol, plan->groups.data(), ng, S(s));
    return e == hipSuccess ? ok() : hip_fail(e, "batch launch");
  }
  // zero fill, then fork: group 0 on the caller's stream, group g > 0 on side stream g - 1
  hipError_t e = pp::launch_batch_zero(plan->op, plan->dvols, plan->nvol, S(s));
  if (e == hipSuccess) e = hipEventRecord(plan->fork, S(s));
  for (int gi = 1; gi < ng && e == hipSuccess; ++gi) {
    const pp_batch_group& g = plan->groups[gi];
    hipStream_t ss = plan->side[gi - 1];
    e = hipStreamWaitEvent(ss, plan->fork, 0);
    if (e == hipSuccess)
      e = pp::launch_batch_group(plan->op, g.dp, g.an != 0, plan->dvols + g.first, g.nvol, g.split, ss);
    if (e == hipSuccess) e = hipEventRecord(plan->join[gi - 1], ss);
  }
  if (e == hipSuccess) {
    const pp_batch_group& g = plan->groups[0];
    e = pp::launch_batch_group(plan->op, g.dp, g.an != 0, plan->dvols + g.first, g.nvol, g.split, S(s));
  }
  for (int gi = 1; gi < ng && e == hipSuccess; ++gi) e = hipStreamWaitEvent(S(s), plan->join[gi - 1], 0);
  return e == hipSuccess ? ok() : hip_fail(e, "batch launch");
}

int pp_batch_plan_destroy(pp_batch_plan* plan) {
  if (!plan) return ok();
  for (hipStream_t ss : plan->side) (void)hipStreamDestroy(ss);
  for (hipEvent_t ev : plan->join) (void)hipEventDestroy(ev);
  if (plan->fork) (void)hipEventDestroy(plan->fork);
  hipError_t e = hipFree(plan->dvols);
  delete plan;
  return e == hipSuccess ? ok() : hip_fail(e, "hipFree");
}

// ---- a batch without allocation ----------------------------------------------------
// pp_batch_run_static keeps the descriptor table in a fixed device array instead of a
// plan's hipMalloc'ed one; the descriptors travel as kernel arguments of a small upload
// kernel (no host staging buffer, nothing whose lifetime outlives the call).
}  // extern "C"

namespace {

constexpr int kStaticVols = 256;  // volumes per pp_batch_run_static call
constexpr int kUploadVols = 8;    // descriptors per upload launch
constexpr int kMaxDevices = 64;

// The table, one copy per device (held until the process ends). Host code takes each
// device's address from hipGetSymbolAddress and never uses this name as an address.
static __device__ __align__(256) pp::BatchVol scratch[kStaticVols];

struct VolChunk {
  pp::BatchVol v[kUploadVols];
};
static_assert(sizeof(VolChunk) <= 4000, "upload chunk must fit the kernel-argument limit");
static_assert(sizeof(pp::BatchVol) % 4 == 0, "descriptors are copied as 32-bit words");

__global__ void __launch_bounds__(256) k_upload_vols(pp::BatchVol* dst, VolChunk src, int n) {
  const unsigned* w = reinterpret_cast<const unsigned*>(&src);
  unsigned* d = reinterpret_cast<unsigned*>(dst);
  const int words = n * (int)(sizeof(pp::BatchVol) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) d[i] = w[i];
}

std::mutex g_static_mu;  // guards the two arrays below (host side of the call only)
pp::BatchVol* g_static_ptr[kMaxDevices];
pp::BatchVol g_static_host[kStaticVols];

}  // namespace

extern "C" {

int pp_batch_run_static(int op, const pp_volume* vols, int nvol, void* stream) {
  if (!vols || nvol < 1) return fail(PP_EINVAL, "need at least one volume");
  if (nvol > kStaticVols)
    return fail(PP_ERANGE, "pp_batch_run_static takes at most %d volumes (got %d); use a batch plan", kStaticVols,
                nvol);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(PP_EDEVICE, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {  // first call on this device, which is current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_static_ptr[dev] = static_cast<pp::BatchVol*>(p);
  }
  pp::BatchVol* d = g_static_ptr[dev];
  pp_batch_group g[kDegGroups];
  int ng = 0;
  const int st = build_batch(op, vols, nvol, g_static_host, g, &ng);
  if (st != PP_OK) return st;
  for (int i = 0; i < nvol; i += kUploadVols) {
    const int n = std::min(kUploadVols, nvol - i);
    VolChunk c{};
    std::memcpy(c.v, g_static_host + i, sizeof(pp::BatchVol) * (size_t)n);
    hipLaunchKernelGGL(k_upload_vols, dim3(1), dim3(256), 0, S(stream), d + i, c, n);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "descriptor upload");
  }
  e = run_batch(static_cast<pp::Op>(op), d, nvol, g, ng, S(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "batch launch");
}

// ---- HIP graphs --------------------------------------------------------------------
int pp_graph_begin(void* s) {
  if (!s) return fail(PP_EINVAL, "graph capture needs a non-default stream");
  hipError_t e = hipStreamBeginCapture(S(s), hipStreamCaptureModeThreadLocal);
  return e == hipSuccess ? ok() : hip_fail(e, "hipStreamBeginCapture");
}

int pp_graph_end(void* s, void** exec) {
  if (!s || !exec) return fail(PP_EINVAL, "null stream or output");
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(S(s), &g);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  *exec = x;
  return ok();
}

int pp_graph_launch(void* exec, void* s) {
  if (!exec) return fail(PP_EINVAL, "null graph");
  hipError_t e = hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(exec), S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "hipGraphLaunch");
}

int pp_graph_destroy(void* exec) {
  if (!exec) return ok();
  hipError_t e = hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(exec));
  return e == hipSuccess ? ok() : hip_fail(e, "hipGraphExecDestroy");
}

int pp_malloc(void** p, size_t bytes) {
  if (!p) return fail(PP_EINVAL, "null pointer");
  hipError_t e = hipMalloc(p, bytes);
  return e == hipSuccess ? ok() : fail(PP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
}
int pp_free(void* p) {
  hipError_t e = hipFree(p);
  return e == hipSuccess ? ok() : hip_fail(e, "hipFree");
}
int pp_memcpy_h2d(void* dst, const void* src, size_t bytes, void* s) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s));
  if (e == hipSuccess) e = hipStreamSynchronize(S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "memcpy h2d");
}
int pp_memcpy_d2h(void* dst, const void* src, size_t bytes, void* s) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s));
  if (e == hipSuccess) e = hipStreamSynchronize(S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "memcpy d2h");
}
int pp_memset_zero(float* p, size_t count, void* s) {
  if (int r = check_mem({{p, "p"}})) return r;
  hipError_t e = pp::launch_fill(p, (long long)count, 0.0f, S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "zero fill");
}
int pp_axpy(float* y, const float* x, float a, size_t count, void* s) {
  if (!y || !x) return fail(PP_EINVAL, "null array");
  if (int r = check_mem({{y, "y"}, {x, "x"}})) return r;
  hipError_t e = pp::launch_axpy(y, x, a, (long long)count, S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "axpy");
}
int pp_lds_atomic_calibrate(unsigned long long* out, int blocks, int iters, void* s) {
  if (!out) return fail(PP_EINVAL, "null array");
  if (blocks < 1 || blocks > 65535 || iters < 1) return fail(PP_EINVAL, "blocks must be 1..65535, iters >= 1");
  hipError_t e = pp::launch_lds_add_rate(out, blocks, iters, S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "lds atomic calibration");
}
int pp_stream_sync(void* s) {
  hipError_t e = hipStreamSynchronize(S(s));
  return e == hipSuccess ? ok() : hip_fail(e, "hipStreamSynchronize");
}
int pp_device_sync(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? ok() : hip_fail(e, "hipDeviceSynchronize");
}

int pp_set_option(int key, int value) {
  switch (key) {
    case PP_OPT_TILING:
      if (value < 0 || value > 2) return fail(PP_EINVAL, "PP_OPT_TILING takes 0, 1 or 2, not %d", value);
      g_tiling.store(value);
      return ok();
    case PP_OPT_TILE_SHAPE:
      if (value < 0 || value > 2) return fail(PP_EINVAL, "tile shape %d not in 0..2", value);
      g_shape.store(value);
      return ok();
    case PP_OPT_USE_TUNED:
      g_tuned.store(value != 0);
      return ok();
    case PP_OPT_BOX_CAP:
      if (value != 0 && (value < 256 || value > kMaxCap))
        return fail(PP_EINVAL, "box budget %d floats not 0 or in 256..%d", value, kMaxCap);
      g_cap.store(value);
      return ok();
    default:
      return fail(PP_EINVAL, "unknown option %d", key);
  }
}

int pp_get_option(int key) {
  switch (key) {
    case PP_OPT_TILING: return g_tiling.load();
    case PP_OPT_TILE_SHAPE: return g_shape.load();
    case PP_OPT_USE_TUNED: return g_tuned.load();
    case PP_OPT_BOX_CAP: return g_cap.load();
    default: return fail(PP_EINVAL, "unknown option %d", key);
  }
}

// ---- autotuner -----------------------------------------------------------------
int pp_tune(int op, const pp_desc* d, float* out, float* out2, double* loss, const float* phi, const float* f,
            const float* aux, void* s, int reps, int* variant) {
  Geo g;
  t_ncand = 0;
  if (int r = check_desc(d, g)) return r;
  if (!uniform_bnd(d))
    return fail(PP_EINVAL, "pp_tune times descriptors whose channels share one boundary triple "
                           "(other mixes run fixed kernels and are not looked up in the table)");
  if (reps < 1) reps = 3;
  auto call = [&]() -> int {
    switch (op) {
      case PP_OP_PULL: return pp_pull(d, out, phi, f, s);
      case PP_OP_PULL_GRAD: return pp_pull_grad(d, out, phi, f, s);
      case PP_OP_PUSH: return pp_push(d, out, phi, f, 0, s);
      case PP_OP_PUSH_GRAD: return pp_push_grad(d, out, phi, f, 0, s);
      case PP_OP_PULL_BACKWARD: return pp_pull_backward(d, out, phi, f, aux, s);
      case PP_OP_OBJECTIVE: return pp_objective_grad(d, out, out2, loss, phi, f, aux, 0, s);
      case PP_OP_PULL_VJP: return pp_pull_vjp(d, out, out2, phi, f, aux, 0, s);
      default: return fail(PP_EINVAL, "op %d has a single kernel variant", op);
    }
  };
  // candidates: per sample; each tile shape at the default budget and at each other budget
  // (anisotropic degrees run on the tiled kernels of the largest)
  const int dmax = std::max(d->deg[0], std::max(d->deg[1], d->deg[2]));
  const int dflt = dmax >= 1 ? pp::default_cap(static_cast<pp::Op>(op), dmax + 1) : 0;
  Choice cand[kMaxCandidates];
  int nc = 0;
  cand[nc++] = Choice{0, 0};
  for (int v = 1; v < kVariants; ++v) {
    cand[nc++] = Choice{v, 0};
    if (dmax < 1) continue;  // no box at degree 0
    for (int c : kTuneCaps)
      if (c != dflt) cand[nc++] = Choice{v, c};
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return fail(PP_EDEVICE, "hipEventCreate");
  float best_ms = 0.f;
  int best = -1, rc = PP_OK;
  for (int k = 0; k < nc && rc == PP_OK; ++k) {
    t_force = cand[k];
    rc = call();  // warm-up (all variants are prebuilt: nothing compiles)
    if (rc != PP_OK) break;
    if (hipEventRecord(e0, S(s)) != hipSuccess) {
      rc = fail(PP_EDEVICE, "hipEventRecord");
      break;
    }
    for (int i = 0; i < reps && rc == PP_OK; ++i) rc = call();
    if (rc == PP_OK && (hipEventRecord(e1, S(s)) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
      rc = fail(PP_EDEVICE, "hipEventRecord/hipEventSynchronize");
    float ms = 0.f;
    if (rc == PP_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
      t_cand[t_ncand] = cand[k];
      t_cand_ms[t_ncand++] = ms / reps;
      if (best < 0 || ms < best_ms) {
        best_ms = ms;
        best = k;
      }
    }
  }
  t_force = Choice{-1, 0};
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PP_OK) return rc;
  if (best < 0) return fail(PP_EDEVICE, "no candidate could be timed");
  const int internal = op;  // PP_OP_* codes equal the internal op numbers
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[{internal, dp_key(d->deg[0] + 1, d->deg[1] + 1, d->deg[2] + 1), size_class(g.n1)}] = cand[best];
  }
  if (variant) *variant = cand[best].variant;
  return ok();
}

int pp_tune_results(int max, int* variants, int* caps, float* ms) {
  const int n = std::min(max, t_ncand);
  for (int i = 0; i < n; ++i) {
    if (variants) variants[i] = t_cand[i].variant;
    if (caps) caps[i] = t_cand[i].cap;
    if (ms) ms[i] = t_cand_ms[i];
  }
  t_status = PP_OK;
  return t_ncand;
}

int pp_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
  return ok();
}

int pp_tune_save(const char* path) {
  if (!path) return fail(PP_EINVAL, "null path");
  FILE* fp = std::fopen(path, "w");
  if (!fp) return fail(PP_EINVAL, "cannot write %s", path);
  std::fprintf(fp, "# pushpull_hip tuning table: op dp log2(voxels) variant box_budget (dp: degree + 1; a degree "
                  "triple: 100 + 36 dpx + 6 dpy + dpz)\n");
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const auto& kv : g_tune)
      std::fprintf(fp, "%d %d %d %d %d\n", std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                   kv.second.variant, kv.second.cap);
  }
  std::fclose(fp);
  return ok();
}

int pp_tune_load(const char* path) {
  if (!path) return fail(PP_EINVAL, "null path");
  FILE* fp = std::fopen(path, "r");
  if (!fp) return fail(PP_EINVAL, "cannot read %s", path);
  char line[256];
  int n = 0;
  std::map<std::tuple<int, int, int>, Choice> t;
  while (std::fgets(line, sizeof line, fp)) {
    if (line[0] == '#') continue;
    int op, dp, sc, v, cap = 0;
    const int got = std::sscanf(line, "%d %d %d %d %d", &op, &dp, &sc, &v, &cap);
    if (got < 4 || v < 0 || v >= kVariants || (cap != 0 && (cap < 256 || cap > kMaxCap))) {
      std::fclose(fp);
      return fail(PP_EINVAL, "%s: malformed line %d", path, n + 1);
    }
    t[{op, dp, sc}] = Choice{v, cap};
    ++n;
  }
  std::fclose(fp);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  for (const auto& kv : t) g_tune[kv.first] = kv.second;
  return ok();
}

const char* pp_last_error(void) { return t_err; }
int pp_last_status(void) { return t_status; }
const char* pp_version(void) { return "pushpull_hip 0.1.0 (gfx950)"; }

}  // extern "C"
