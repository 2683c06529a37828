/*
 * pp_oracle.c -- CPU oracle for the B-spline pull/push path of PushPull.jl.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker. The HIP library never links it and
 * never falls back to it.
 *
 * Two instantiations of pp_oracle_impl.h:
 *   ppo_{pull,pullg,pullh,push,pushg}_f32  FP32 restatement of artifacts/C/pushpull.c
 *                                          (bit-exact against the reference compiled
 *                                          from source; see tests/test_oracle_golden.py)
 *   ppo_{pull,pullg,pullh,push,pushg}_f64  FP64 oracle: x = phi - 1.0f, the validity
 *                                          mask and the knot ceilf(x - dp/2) are taken in
 *                                          FP32 as the reference takes them; weights,
 *                                          gathers and sums are in double.
 * Signatures follow artifacts/C/pushpull.h:3-14 (plus pullh, pushpull.c:70): the
 * caller passes one (channel, batch) slice; d0 is the source grid for the pull family
 * and the destination grid for the push family; dp = degree + 1 in [1, 5].
 *
 * Deliberate deviation from the reference CPU code: ppo_pushg_* is the true adjoint of
 * ppo_pullg_* (pushpull.cu:145-163). The CPU pushg at pushpull.c:144 tests
 * `!ISFINITE(g[1])` and therefore pushes nothing for finite input.
 */
#include <math.h>
#include <stdint.h>

typedef unsigned long long u64;
#define PPO_MAXD 5

/* Knot of a sample: (int64)ceilf(x - 0.5f*dp)   (bspline.cu:66).
 * The reference converts with a plain C cast, which on x86-64 is cvttss2si: a value
 * outside the int64 range yields INT64_MIN. That is reproduced here explicitly so
 * the result does not depend on the compiler's handling of the undefined cast. */
static long long ppo_knot(float x, int dp)
{
    float c = ceilf(x - 0.5f * (float)dp);
    if (!(c >= -9223372036854775808.0f && c < 9223372036854775808.0f)) return INT64_MIN;
    return (long long)c;
}

/* Boundary folding of index i onto [0, d) with a sign   (patch.cu:5-59):
 *   code 0: circulant;  code 2: half-sample antisymmetric (sign -1 on mirrored copies);
 *   any other code: half-sample symmetric reflection. */
static void ppo_fold(int code, long long i, u64 d, u64 *b, int *s)
{
    if (code == 0) {
        long long r = i % (long long)d;
        if (r < 0) r += (long long)d;
        *b = (u64)r;
        *s = 1;
    } else {
        long long p = 2 * (long long)d;
        long long m = i % p;
        if (m < 0) m += p;
        if (m < (long long)d) {
            *b = (u64)m;
            *s = 1;
        } else {
            *b = (u64)(p - 1 - m);
            *s = (code == 2) ? -1 : 1;
        }
    }
}

/* VOXOK (pushpull.c:6-9): finite coordinates and, unless extrapolating, inside
 * [-0.01, d-0.99] on every axis, all in FP32. */
static int ppo_voxok(const float *x, const u64 *d0, int ext)
{
    int a;
    for (a = 0; a < 3; a++)
        if (!isfinite(x[a])) return 0;
    if (ext != 0) return 1;
    for (a = 0; a < 3; a++)
        if (!(x[a] >= -0.01f && x[a] <= (float)d0[a] - 0.99f)) return 0;
    return 1;
}

/* exported helpers so tests can pin the index work directly */
long long ppo_knot_export(float x, int dp) { return ppo_knot(x, dp); }
void ppo_fold_export(int code, long long i, u64 d, u64 *b, int *s) { ppo_fold(code, i, d, b, s); }

#define REAL float
#define SFX _f32
#define PPO_IS_DOUBLE 0
#include "pp_oracle_impl.h"
#undef REAL
#undef SFX
#undef PPO_IS_DOUBLE

#define REAL double
#define SFX _f64
#define PPO_IS_DOUBLE 1
#include "pp_oracle_impl.h"
#undef REAL
#undef SFX
#undef PPO_IS_DOUBLE

/* B-spline weight vectors, exported for the closed-form tests */
void ppo_weights_f64(int dp, double t, double *w, double *dw, double *hw)
{
    wts_f64(dp, t, w);
    dwts_f64(dp, t, dw);
    hwts_f64(dp, t, hw);
}
void ppo_weights_f32(int dp, float t, float *w, float *dw, float *hw)
{
    wts_f32(dp, t, w);
    dwts_f32(dp, t, dw);
    hwts_f32(dp, t, hw);
}
