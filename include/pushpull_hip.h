/*
 * pushpull_hip.h -- C ABI of libpushpull_hip.so, the gfx950 (MI355X) implementation of
 * the B-spline pull/push sampling path of PushPull.jl.
 *
 * Plain C types only: host or device pointers, sizes, int status codes; a HIP stream
 * is passed as `void *` (NULL = the default stream).
 *
 * Two tiers.
 *
 * Tier A: symbol-compatible with the reference's native library (artifacts/C/pushpull.h:3-14
 *   plus pullh, artifacts/C/pushpull.c:70). Host pointers, one (channel, batch) slice per
 *   call, synchronous: upload, launch on the GPU, download. The reference's Julia wrappers
 *   (src/pushpull_cpu.jl:45-49, :100-104, :140-144, :195-199, :251-255) can ccall these
 *   unchanged once `pplib` (src/pushpull_init.jl:4) points at this library. Like the
 *   reference they return nothing; a failure is reported through pp_last_error() and
 *   pp_last_status(), and the output is left filled with NaN.
 *
 * Tier B: device-resident, stream-ordered, re-entrant. Arguments are device pointers
 *   holding whole Julia-layout arrays (all channels and batches), described by a pp_desc
 *   passed by pointer and copied into kernel arguments at launch (no module globals,
 *   unlike the reference CUDA module's __constant__ settings, artifacts/C/pushpull.cu:16-24).
 *   Every function returns 0 on success or a negative PP_E* code, with a message in
 *   pp_last_error() (thread-local).
 *
 * Tier B memory contract: every array (and every pointer in a pp_volume) is
 *   coarse-grained device memory of the current device -- hipMalloc / hipMallocAsync
 *   memory, which is what torch and most allocators hand out. The scatters accumulate with
 *   device-scope float atomics that are exact read-modify-writes only there. Host memory
 *   (pinned or not), managed or unified memory and memory of another device are rejected
 *   with PP_EINVAL (hipPointerGetAttributes, checked on the host at every call).
 *   Fine-grained device memory (hipExtMallocWithFlags + hipDeviceMallocFinegrained) is
 *   outside the contract too but reports the same attributes and is not detected.
 *
 * Array layout (column-major, as Julia): image (X0,Y0,Z0,C,B); deformation phi
 * (X1,Y1,Z1,3,B) holding 1-BASED voxel coordinates into the image grid; pull output
 * (X1,Y1,Z1,C,B); pull_grad output (X1,Y1,Z1,3,C,B); pull_hess output (X1,Y1,Z1,3,3,C,B).
 */
#ifndef PUSHPULL_HIP_H
#define PUSHPULL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* USIZE_t of the reference (artifacts/C/cuheader.h:14,22) */
typedef unsigned long long pp_usize;

/* ---------------------------------------------------------------- Tier A ---- */

/* f1[n1] = pull of f0 (grid d0) at phi (3*n1, 1-based). NaN where the sample is invalid.
 * Replaces artifacts/C/pushpull.c:14 `pull`. dp = degree+1 per axis, in [1,5]. */
void pull(float *f1, const float *phi, const float *f0, const pp_usize *d0, pp_usize n1,
          const int *bnd, const pp_usize *dp, int ext);

/* g1[3*n1] = spatial gradient of the pulled image.  Replaces pushpull.c:37 `pullg`. */
void pullg(float *g1, const float *phi, const float *f0, const pp_usize *d0, pp_usize n1,
           const int *bnd, const pp_usize *dp, int ext);

/* h1[9*n1] = Hessian, column-major 3x3 with symmetric copies. Replaces pushpull.c:70 `pullh`. */
void pullh(float *h1, const float *phi, const float *f0, const pp_usize *d0, pp_usize n1,
           const int *bnd, const pp_usize *dp, int ext);

/* f0 (grid d0) += push of f1[n1] at phi: the adjoint of pull. Accumulates into the caller's
 * f0 like the reference. Replaces pushpull.c:106 `push`. */
void push(float *f0, const float *phi, const float *f1, const pp_usize *d0, pp_usize n1,
          const int *bnd, const pp_usize *dp, int ext);

/* f0 += push of gradients g1[3*n1]: the adjoint of pullg. Replaces pushpull.c:127 `pushg`
 * with the semantics of artifacts/C/pushpull.cu:145-163 (the CPU version's inverted
 * finiteness test at pushpull.c:144 is not reproduced). */
void pushg(float *f0, const float *phi, const float *g1, const pp_usize *d0, pp_usize n1,
           const int *bnd, const pp_usize *dp, int ext);

/* f1 (grid d1) = pull of f0 at x = A*(i,j,k,1), A the 0-based 3x4 affine (column-major,
 * 12 floats). Replaces pushpull.c:176 `pull_affine`; the loop runs over d1 (n1 is not used). */
void pull_affine(float *f1, const float *f0, const pp_usize *d0, pp_usize n1, const pp_usize *d1,
                 const float *Aff, const int *bnd, const pp_usize *dp, int ext);

/* f0 (grid d0) += push of f1 (grid d1) at x = A*(i,j,k,1). Replaces pushpull.c:213 `push_affine`. */
void push_affine(float *f0, const float *f1, const pp_usize *d0, pp_usize n1, const pp_usize *d1,
                 const float *Aff, const int *bnd, const pp_usize *dp, int ext);

/* Sparse regulariser (artifacts/C/sparse_operator.c, built by the reference as its own
 * library, `oplib`, src/pushpull_init.jl:5). d = (X, Y, Z, nchan, hcode); the stencil
 * is the `sparsify` tuple of src/operator.jl:206-296 (offset, length: nchan x nchan,
 * column-major; values, patch_indices: the lists). `indices` is accepted and not read:
 * interior offsets are derived from patch_indices and d (sparsify writes both from the
 * same positions). Results are bit-identical to the reference's CPU code. */

/* u = L v. Like the reference, voxels within (dp-1)/2 of a face ADD to the caller's u and
 * the others overwrite it. Replaces sparse_operator.c:45 `vel2mom`. */
void vel2mom(float *u, const float *v, const pp_usize *d, const int *offset, const int *length,
             const float *values, const int *indices, const int *patch_indices, const pp_usize *dp,
             const int *bnd);

/* One Gauss-Seidel sweep of (L + H) v = g, updating v. h has nchan (d[4] == 1) or
 * nchan*(nchan+1)/2 (d[4] == 2) channels, or is unused (d[4] == 0). Replaces
 * sparse_operator.c:116 `relax`. */
void relax(float *v, const pp_usize *d, const float *g, const float *h, const int *offset,
           const int *length, const float *values, const int *indices, const int *patch_indices,
           const pp_usize *dp, const int *bnd);

/* ---------------------------------------------------------------- Tier B ---- */

enum {
  PP_OK = 0,
  PP_EINVAL = -1,   /* bad argument (shape, degree, boundary code, null pointer) */
  PP_EDEVICE = -2,  /* HIP runtime error */
  PP_ENOMEM = -3,   /* allocation failed */
  PP_ERANGE = -4    /* size beyond the supported range */
};

#define PP_MAX_CHAN_BND 64

/* Settings of one call (src/pushpull_general.jl:15-60) plus the array geometry. */
typedef struct pp_desc {
  int64_t d0[3];   /* image grid (X,Y,Z): sampled by pull, scattered into by push */
  int64_t d1[3];   /* grid of phi (X,Y,Z) */
  int32_t nchan;   /* C */
  int32_t nbatch;  /* B */
  int32_t deg[3];  /* interpolation degree per axis, 0..4 */
  int32_t ext;     /* 0: samples outside [-0.01, d-0.99] are invalid; 1: extrapolate */
  int32_t nbnd;    /* 1 (shared) or nchan (per channel) */
  int32_t bnd[3 * PP_MAX_CHAN_BND]; /* boundary codes per axis (0 circulant, 1 reflect, 2 antisymmetric) */
} pp_desc;

/* Julia: pull(f0, phi, sett)            src/pushpull_cpu.jl:17      out (d1,C,B) */
int pp_pull(const pp_desc *d, float *f1, const float *phi, const float *f0, void *stream);
/* Julia: pull_grad(f0, phi, sett)       src/pushpull_cpu.jl:69      out (d1,3,C,B) */
int pp_pull_grad(const pp_desc *d, float *g1, const float *phi, const float *f0, void *stream);
/* Julia: pull_hess(f0, phi, sett)       src/pushpull_cpu.jl:111     out (d1,3,3,C,B) */
int pp_pull_hess(const pp_desc *d, float *h1, const float *phi, const float *f0, void *stream);
/* Julia: push(f1, phi, d0, sett)        src/pushpull_cpu.jl:165     out (d0,C,B);
 * accumulate=0 zeroes f0 first (in-stream), 1 adds to its contents. */
int pp_push(const pp_desc *d, float *f0, const float *phi, const float *f1, int accumulate, void *stream);
/* Julia: push_grad(g1, phi, d0, sett)   src/pushpull_cpu.jl:220     in (d1,3,C,B) */
int pp_push_grad(const pp_desc *d, float *f0, const float *phi, const float *g1, int accumulate, void *stream);

/* Julia: affine_pull(f0, Aff, d1, sett)  src/pushpull_cpu.jl:268  out (d1,C,B); `aff` is the
 * 0-based 3x4 matrix (12 floats, column-major, host memory) that adjust_affine
 * (src/pushpull_cpu.jl:357-362) produces; d->d1 is the output grid. */
int pp_affine_pull(const pp_desc *d, float *f1, const float *f0, const float *aff, void *stream);
/* Julia: affine_push(f1, Aff, d0, sett)  src/pushpull_cpu.jl:309  out (d0,C,B) */
int pp_affine_push(const pp_desc *d, float *f0, const float *f1, const float *aff, int accumulate, void *stream);

/* Pullback of pull wrt phi (src/pushpull_pb.jl:13-17): dphi(d1,3,B) = sum_c delta_c * pull_grad(f0)_c,
 * NaN where the sample is invalid. Also the phi-pullback of push (src/pushpull_pb.jl:41-51)
 * with (f0, delta) := (upstream gradient, pushed image). */
int pp_pull_backward(const pp_desc *d, float *dphi, const float *phi, const float *f0, const float *delta,
                     void *stream);
/* Pullback of pull_grad wrt phi (src/pushpull_pb.jl:26-35):
 * dphi_b = sum_c sum_a delta(a,c) * pull_hess(f0)(a,b,c); delta is (d1,3,C,B). */
int pp_pull_grad_backward(const pp_desc *d, float *dphi, const float *phi, const float *f0, const float *delta,
                          void *stream);

/* Fused objective and both gradients for L = 0.5*||pull(f0,phi) - g||^2 (no counterpart in
 * the reference; it composes pull, push and pp_pull_backward):
 *   dphi (d1,3,B) = sum_c r_c grad pull(f0)_c (NaN where invalid), r = pull(f0,phi) - g,
 *   df0  (d0,C,B) = push(r) (zeroed first unless accumulate), loss[0] (double, device) =
 *   0.5 * sum of r^2 over finite r (zeroed first unless accumulate). */
int pp_objective_grad(const pp_desc *d, float *dphi, float *df0, double *loss, const float *phi,
                      const float *f0, const float *g, int accumulate, void *stream);
/* pp_objective_grad followed by the descent step phi_next = fmaf(step, dphi, phi), voxel by
 * voxel (no counterpart in the reference: a registration loop's update fused into the
 * gradient pass). phi_next may be phi itself (the update in place); it must not overlap dphi.
 * Rounded exactly as pp_axpy(phi_next, dphi, step) after a copy of phi would round it. One
 * launch where the channels share their boundary codes (or differ only between reflect and
 * antisymmetric, at one degree) and the tiled kernels run; else the objective, then the step. */
int pp_objective_grad_step(const pp_desc *d, float *dphi, float *df0, double *loss, const float *phi,
                           const float *f0, const float *g, float *phi_next, float step, int accumulate,
                           void *stream);

/* Both cotangents of pull in one pass (the rrule of pull, src/pushpull_pb.jl:11-17, whose
 * two thunks call push and pull_grad separately; the reference has no fused entry point):
 *   dphi (d1,3,B) = sum_c delta_c grad pull(f0)_c, NaN where the sample is invalid
 *                   (= pp_pull_backward),
 *   df0  (d0,C,B) = push(delta) (zeroed first unless accumulate; = pp_push).
 * One launch stages each tile of f0 once and reads phi and delta once for both. */
int pp_pull_vjp(const pp_desc *d, float *dphi, float *df0, const float *phi, const float *f0,
                const float *delta, int accumulate, void *stream);

/* ---- push with a count image ------------------------------------------------
 * pp_push_count replaces pushc (pushpull.cu:128-142: push1 with its count at
 * pushpull_dev.cu:140-154) and the push pairs of its callers: invert_def pushes the
 * identity and ones through one deformation (src/multigrid.jl:270-273); restrict pushes an
 * image and ones through one affine (src/multigrid.jl:26-30, via pp_affine_push_count).
 * f0 (d0, nchan, nbatch) receives push(f1) channel by channel with each channel's
 * boundary triple; c0 (d0, 1, nbatch) receives the push of ones with the count triple:
 * triple number nchan of d->bnd when d->nbnd == nchan + 1, else the first triple (as in
 * pushc, whose count shares the image's boundary, signs included). f1 == NULL pushes the
 * identity instead: channel c holds the 1-based index along axis c of the phi grid (nchan
 * must be 3). Invalid samples and non-finite values are skipped as in push; accumulate = 0
 * zeroes f0 and c0 first. One launch when every triple folds alike (equal, or differing
 * only between 1 and 2) and the degree is isotropic 1-3; otherwise one push per channel
 * and one for the count. */
int pp_push_count(const pp_desc *d, float *f0, float *c0, const float *phi, const float *f1, int accumulate,
                  void *stream);
int pp_affine_push_count(const pp_desc *d, float *f0, float *c0, const float *f1, const float *aff, int accumulate,
                         void *stream);

/* ---- compositions (src/compositions.jl) -------------------------------------
 * pp_compose replaces compose!/compose (src/compositions.jl:9-39): psi = phi o theta for
 * a displacement phi (3 channels on grid d0) sampled at the 1-based deformation theta
 * (grid d1): channel c is pulled with degree 1, extrapolation, and boundary 1 on every axis
 * except 2 on axis c (the "sliding" matrix of compositions.jl:33), all three channels in
 * one launch. psi has theta's size; nbatch volumes lie back to back. psi must not overlap
 * phi or theta.
 * pp_exp replaces Exp (src/compositions.jl:41-49): scaling and squaring of a velocity v
 * (3 channels, grid dim): u = v / 2^nsq, then nsq times u = compose(u, Id + u), then
 * out = u + Id (1-based). The reference squares 8 times. Id + u is formed inside the kernel
 * with the reference's float rounding. work is a caller buffer of v's size; out, v and work
 * must not overlap. Both are stream-ordered and graph-capturable. */
int pp_compose(float *psi, const float *phi, const float *theta, const long long d0[3], const long long d1[3],
               int nbatch, void *stream);
int pp_exp(float *out, const float *v, float *work, const long long dim[3], int nbatch, int nsq, void *stream);

/* ---- sparse regulariser (src/operator_sparse.jl, src/operator.jl) ------------
 * A stencil L in the `sparsify` list format (src/operator.jl:206-296) is uploaded once
 * for volumes of size d; the calls then take device arrays (X, Y, Z, nchan), column-major,
 * and a host array bnd of 3*nchan boundary codes (channel c's triple at 3c). nchan <= 8,
 * stride odd and <= 5 per axis. Results are bit-identical to artifacts/C/sparse_operator.c.
 * The kernel object is tied to the device current at creation. */
typedef struct pp_sparse_kernel pp_sparse_kernel;

/* offset/length: nchan*nchan (pair (i, j) at i + nchan*j, i the output channel); the list
 * of a pair is values/patch_indices[offset .. offset+length), values[offset-1] its sum of
 * non-centre values. Host arrays, copied. Replaces the kernel tuple passed to every call
 * of src/operator_sparse.jl:16-62. */
int pp_sparse_create(pp_sparse_kernel **kernel, const int64_t d[3], int nchan, const int32_t stride[3],
                     const int32_t *offset, const int32_t *length, const float *values,
                     const int32_t *patch_indices, int64_t nvalues);
int pp_sparse_destroy(pp_sparse_kernel *kernel);

/* u = L v (src/operator_sparse.jl:6-40). accumulate_edges = 1 reproduces vel2mom!'s
 * effect on a non-zero u exactly (voxels within (dp-1)/2 of a face add to u); 0 overwrites
 * u everywhere (vel2mom). u and v must be different arrays. */
int pp_vel2mom(const pp_sparse_kernel *kernel, float *u, const float *v, const int32_t *bnd,
               int accumulate_edges, void *stream);

/* nit Gauss-Seidel sweeps of (L + H) v = g, updating v in place (src/operator_sparse.jl:43-62).
 * hcode 0: no H (h may be NULL); 1: h holds nchan diagonal channels; 2: nchan*(nchan+1)/2
 * channels, the diagonal first and then the pairs (j, i > j) in order. Boundary codes must
 * be 0, 1 or 2 (src/operator_sparse.jl:82-84). Each sweep runs the reference's colour
 * order as one launch per colour of each of its seven regions. */
int pp_relax(const pp_sparse_kernel *kernel, float *v, const float *g, const float *h, int hcode,
             const int32_t *bnd, int nit, void *stream);

/* ---- batches of uneven volumes --------------------------------------------
 * No counterpart in the reference, whose wrappers loop over same-size (channel, batch)
 * slices with one native call each (src/pushpull_cpu.jl:42-51, ext/pushpull_gpu.jl:37-48).
 * A plan describes many volumes of different sizes, channel counts, degrees (a triple per
 * volume) and boundary codes (a triple per volume or per channel); running it is one
 * persistent launch over all the tiles of the volumes that share a degree group (their
 * largest degree, and whether their degrees differ between axes; plus one zeroing launch
 * for the accumulating ops), on kernels only, so it can be captured into a HIP graph. */
#define PP_BATCH_CHAN_BND 10 /* channels of a volume that may carry boundary codes of their own */
enum {
  PP_OP_PULL = 0,          /* out = pull(f, phi)                                   */
  PP_OP_PULL_GRAD = 1,     /* out = pull_grad(f, phi)                              */
  PP_OP_PUSH = 3,          /* out = push(f, phi) (zeroed first)                    */
  PP_OP_PUSH_GRAD = 4,     /* out = push_grad(f, phi) (zeroed first)               */
  PP_OP_PULL_BACKWARD = 5, /* out = dphi = sum_c aux_c grad pull(f)_c              */
  PP_OP_OBJECTIVE = 7,     /* out = dphi, out2 = dL/df (zeroed first), loss[0] = L */
  PP_OP_PULL_VJP = 8       /* out = dphi, out2 = push(aux) (zeroed first)          */
};

typedef struct pp_volume {
  const float *f;    /* image (d0,C) for the pull family; pushed values (d1,C) or (d1,3,C) for push */
  const float *phi;  /* (d1,3), 1-based */
  const float *aux;  /* PULL_BACKWARD, PULL_VJP: upstream gradient (d1,C); OBJECTIVE: target g (d1,C) */
  float *out;        /* PULL: (d1,C); PUSH/PUSH_GRAD: (d0,C); PULL_BACKWARD/OBJECTIVE/PULL_VJP: dphi (d1,3) */
  float *out2;       /* OBJECTIVE: dL/df (d0,C); PULL_VJP: push(aux) (d0,C) (both zeroed by the run) */
  double *loss;      /* OBJECTIVE: 0.5*sum r^2 of this volume (device, one double) */
  int64_t d0[3];     /* image grid (X,Y,Z) */
  int64_t d1[3];     /* phi grid (X,Y,Z) */
  int32_t nchan;
  int32_t deg[3];    /* degree 0..4 per axis (X,Y,Z), as pp_desc */
  int32_t ext;
  int32_t bnd[3];    /* shared by the volume's channels (nbnd 0 or 1) */
  /* nbnd = nchan (<= PP_BATCH_CHAN_BND): chan_bnd[3c + axis] is channel c's code instead
   * (src/pushpull_cpu.jl:43 takes a triple per channel). Along an axis the channels' codes
   * must be equal or all in {1, 2} (reflect, antisymmetric: the same fold, a sign apart);
   * other mixes take single calls. */
  int32_t nbnd;
  int32_t chan_bnd[3 * PP_BATCH_CHAN_BND];
  /* OBJECTIVE only, optional (no counterpart in the reference: a registration loop's descent
   * step, phi <- phi - alpha dL/dphi, fused into the gradient pass): when set, phi_next
   * receives fmaf(step, dphi, phi) voxel by voxel, rounded as pp_axpy(phi_next, dphi, step)
   * would; it may be phi itself (each voxel's phi is read only by the tile computing its
   * dphi). Must be NULL for the other ops. */
  float *phi_next;
  float step;
} pp_volume;

typedef struct pp_batch_plan pp_batch_plan;
/* Validates the volumes and uploads their descriptors (allocates device memory once). */
int pp_batch_plan_create(int op, const pp_volume *vols, int nvol, pp_batch_plan **plan);
/* Stream-ordered; allocates nothing; capturable. */
int pp_batch_run(const pp_batch_plan *plan, void *stream);
int pp_batch_plan_destroy(pp_batch_plan *plan);
/* The same batch in one call that allocates nothing: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Runs op over vols exactly as a plan of them
 * would (same kernels, same results), taking the descriptor table from a fixed per-device
 * array of at most 256 volumes (nvol > 256: PP_ERANGE) and uploading it through kernel
 * arguments on the stream. Every call on one device shares that array and leaves the last
 * call's table in it: calls on one stream are ordered, but calls that overlap on one GPU
 * (other streams, other threads without a stream order between them) race. The pointers
 * in vols must stay valid until the stream has run the call. */
int pp_batch_run_static(int op, const pp_volume *vols, int nvol, void *stream);

/* ---- HIP graph capture of any sequence of the calls above (stream must not be NULL) ---- */
int pp_graph_begin(void *stream);
int pp_graph_end(void *stream, void **graph_exec);
int pp_graph_launch(void *graph_exec, void *stream);
int pp_graph_destroy(void *graph_exec);

/* ---- device memory helpers, so a Julia caller needs no GPU package ---------- */
int pp_malloc(void **p, size_t bytes);
int pp_free(void *p);
int pp_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int pp_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int pp_memset_zero(float *p, size_t count, void *stream);
/* y[i] += a * x[i] (stream-ordered kernel; e.g. the deformation update of a matching loop) */
int pp_axpy(float *y, const float *x, float a, size_t count, void *stream);
/* Calibration of the scatters' on-chip bound (no counterpart in the reference): `blocks`
 * workgroups of 256 lanes each add the 64 taps of `iters` cubic samples into an LDS box
 * shaped like a tile's, one integer LDS atomic per tap on 64-bit cells as the
 * scatters use (blocks * 256 * iters * 64 adds); out (blocks * 256 words, device)
 * receives checksums. Time it on `stream`. */
int pp_lds_atomic_calibrate(unsigned long long *out, int blocks, int iters, void *stream);
int pp_stream_sync(void *stream);
int pp_device_sync(void);

/* ---- kernel selection (no counterpart in the reference) -------------------- */
enum {
  PP_OPT_TILING = 1,     /* 1 (default): automatic -- the autotuner's choice where it has one,
                            else LDS-tiled kernels (one work-item per sample for degree 0 and
                            for trilinear gathers over a one-row grid, as Tier A passes);
                            2: LDS-tiled kernels wherever they exist (pull, pull_grad,
                            pull_hess, push, push_grad, the pullbacks, the pull VJP, the fused
                            objective; degrees 1-4, a degree per axis running on the 16x8x8
                            kernels of the largest); 0: one work-item per sample */
  PP_OPT_TILE_SHAPE = 2, /* tile of the phi grid: 0 = 16x8x8 (default), 1 = 32x8x4, 2 = 16x16x4 */
  PP_OPT_USE_TUNED = 3,  /* 1 (default): prefer the autotuner's choice when it has one */
  PP_OPT_BOX_CAP = 4     /* LDS box budget of the tiled kernels in floats (256..16384); 0
                            (default): a per-op, per-degree default. A box beyond the budget
                            is clipped to it and the samples it leaves out run per sample
                            through global memory; a larger budget leaves fewer workgroups per
                            CU */
};
int pp_set_option(int key, int value);
int pp_get_option(int key);

/* ---- autotuner (no counterpart in the reference) ----------------------------
 * Times the kernel candidates of `op` (PP_OP_PULL, _PULL_GRAD, _PUSH, _PUSH_GRAD,
 * _PULL_BACKWARD, _OBJECTIVE, _PULL_VJP) on the caller's buffers -- one warm-up and `reps`
 * timed calls each, HIP events on `stream` -- and records the fastest for (op, degree,
 * log2 of the phi-grid size); later calls with that key use it. Candidates: one work-item
 * per sample, and each tile shape at the op's default LDS box budget and at the budgets
 * that let 6, 5, 4 and 3 workgroups share a CU (largest degree >= 1; a degree per axis
 * runs on the 16x8x8 tiles whatever the shape asked for). The buffers are
 * those of the op's Tier B function (out2 for the objective and the VJP, loss for the
 * objective, aux for the pullback, the objective and the VJP). The descriptor's channels
 * must share one boundary triple (PP_EINVAL otherwise: other mixes run fixed kernels).
 * *variant receives 0 (one work-item per sample) or 1+k (tiled, tile shape k). */
int pp_tune(int op, const pp_desc *d, float *out, float *out2, double *loss, const float *phi, const float *f,
            const float *aux, void *stream, int reps, int *variant);
/* The candidates of this thread's last pp_tune, in the order timed: variant, box budget in
 * floats (0: the default) and mean ms per call, up to `max` of each (arrays may be NULL).
 * Returns how many were timed. */
int pp_tune_results(int max, int *variants, int *caps, float *ms);
int pp_tune_save(const char *path);  /* text table: op dp log2(voxels) variant box_budget */
int pp_tune_load(const char *path);  /* also reads 4-column tables (budget: default) */
int pp_tune_clear(void);

/* ---- diagnostics ---------------------------------------------------------- */
const char *pp_last_error(void);  /* message of the last failure on this thread */
int pp_last_status(void);          /* status of the last call on this thread */
const char *pp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PUSHPULL_HIP_H */
