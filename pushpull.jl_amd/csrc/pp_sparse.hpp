// pp_sparse.hpp -- host-side view of the sparse-regulariser kernels (internal to
// libpushpull_hip): u = L v (vel2mom) and the Gauss-Seidel update of (L + H) v = g (relax)
// for a sparse stencil L in the list format of src/operator.jl:206-296 (`sparsify`).
#pragma once
#include <hip/hip_runtime.h>

#include "pp_kernels.hpp"

namespace pp {

constexpr int kSpMaxChan = 8;  // MAXN of artifacts/C/sparse_operator_dev.cu:8
constexpr int kSpMaxBox = 5;   // largest stencil box side (src/operator_sparse.jl:66,103-104)

// One (output i, input j) list. w0 is the weight of its first entry (the centre, which
// make_A reads, sparse_operator_dev.cu:63) and c0w = w0 + (sum of the list's other
// weights), rounded as one float add (sp_conv's last term, sparse_operator_dev.cu:18).
struct SpPair {
  float w0, c0w;
  int start, len;  // the list's entries in the entry table
};
// The entries of all lists, pair (i, j) = (output, input) at p = i + nc * j, in increasing
// p: the order in which vel2mom and relax consume them (j outer, i inner). goff = nd * j +
// the position's linear offset from the centre voxel (interior voxels read v[m + goff]);
// code packs the box offsets (dx, dy, dz) as bytes biased by 16, then the flags below, the
// output channel i (bits 26-28) and the input channel j (bits 29-31). A list's first entry
// is the value read as f0 and carries c0w in w; the others carry their weights.
struct SpEntry {
  long long goff;
  float w;
  int code;
};
constexpr int kSpFirst = 1 << 24;
constexpr int kSpLast = 1 << 25;

struct SpArgs {
  const PP_GLOBAL SpPair* pairs;  // nc * nc, pair (i, j) at i + nc * j
  const PP_GLOBAL SpEntry* ent;   // nent entries
  int nent;
  PP_GLOBAL float* v;             // vel2mom: input; relax: the estimate, updated in place
  PP_GLOBAL float* u;             // vel2mom: output
  const PP_GLOBAL float* g;       // relax: right-hand side (nc channels)
  const PP_GLOBAL float* h;       // relax: H (nc or nc (nc + 1) / 2 channels), or null
  long long nd;                   // voxels per channel
  int d[3];
  int nc;
  int hcode;                      // 0 no H, 1 diagonal, 2 full symmetric
  int accum;                      // vel2mom: edge voxels add to u rather than overwrite
  int bnd[3 * kSpMaxChan];        // boundary codes of input channel j at 3 j
};

// The voxels start + step * (ix, iy, iz), 0 <= i* < cnt*, numbered with x fastest.
// rx, ry: 1/cnt[0], 1/cnt[1] in double (exact-quotient decode with one correction step).
// split/key (relax under a circulant fold whose period is not a multiple of the colour
// stride): bit 0 / bit 1 of split select the x / y test "within r of the upper end"; a
// voxel runs only when its test bits equal key.
struct SpLattice {
  int start[3], step[3], cnt[3];
  long long n;
  double rx, ry;
  int split, key;
  int r[3];
};

enum SpOp : int { SP_VEL2MOM = 0, SP_RELAX = 1 };

// edge: fold every stencil read through the boundary codes (voxels within r of a face);
// otherwise the lattice lies in the interior and reads use the entries' linear offsets.
hipError_t launch_sparse(SpOp op, const SpArgs& a, const SpLattice& l, bool edge, hipStream_t stream);

// vel2mom over the whole volume in one LDS-tiled launch, for stencils of radius 1 <= r <= 2
// and nc <= 3; `interior` is the reference's interior block (step 1), outside which u
// accumulates when a.accum is set. hipErrorNotSupported otherwise (then launch_sparse
// covers the seven regions).
hipError_t launch_vel2mom_tiled(const SpArgs& a, const SpLattice& interior, int r, hipStream_t stream);

// Error state of the C ABI (pp_last_error / pp_last_status), shared by its source files.
int abi_fail(int code, const char* msg);
int abi_ok();

}  // namespace pp
