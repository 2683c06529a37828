/*
 * Generic body of the CPU oracle, instantiated twice by pp_oracle.c:
 *   REAL=float  SFX=_f32  -> the FP32 restatement (same per-operation order as the
 *                            reference C, so it is bit-exact against it)
 *   REAL=double SFX=_f64  -> the FP64 oracle (mask and knot decided in FP32 exactly
 *                            as the reference decides them; all arithmetic in double)
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product path links or calls this code.
 *
 * Reference files followed (paths relative to the PushPull.jl tree):
 *   artifacts/C/bspline.cu:16-244   B-spline weights and their 1st/2nd derivatives
 *   artifacts/C/patch.cu:5-139      boundary folding, gather and scatter of a patch
 *   artifacts/C/pushpull_dev.cu:7-179  patch contractions (value / gradient / Hessian),
 *                                   push1 / pushg1
 *   artifacts/C/pushpull.c:6-152    per-voxel loops, VOXOK validity, NaN rules
 *   artifacts/C/pushpull.cu:145-163 push_grad semantics (the CPU pushg at
 *                                   pushpull.c:144 inverts a finiteness test and is
 *                                   not followed; see pp_oracle.c)
 */

#define PPO_CAT2(a, b) a##b
#define PPO_CAT(a, b) PPO_CAT2(a, b)
#define FN(name) PPO_CAT(name, SFX)

/* ---- B-spline kernels (bspline.cu:16-59, 94-139, 186-211) ------------------ */
/* The FP32 instance must reproduce the reference's rounding, so each polynomial
 * keeps the reference's operation order. */

static REAL FN(bs2)(REAL u)
{
    u = (REAL)fabs(u);
    if (u < (REAL)0.5) return (REAL)0.75 - u * u;
    if (u < (REAL)1.5) { u = (REAL)1.5 - u; return (REAL)0.5 * u * u; }
    return 0;
}

static REAL FN(bs3)(REAL u)
{
    u = (REAL)fabs(u);
    if (u < 1) return u * u * (u - (REAL)2.0) * (REAL)0.5 + (REAL)2.0 / (REAL)3.0;
    if (u < 2) { u = (REAL)2.0 - u; return u * u * u * ((REAL)1.0 / (REAL)6.0); }
    return 0;
}

static REAL FN(bs4)(REAL u)
{
    u = (REAL)fabs(u);
    if (u < (REAL)0.5) {
        u *= u;
        return u * (u * (REAL)0.25 - (REAL)0.625) + (REAL)115.0 / (REAL)192.0;
    }
    if (u < (REAL)1.5)
        return u * (u * (u * ((REAL)5.0 / (REAL)6.0 - u * ((REAL)1.0 / (REAL)6.0)) - (REAL)1.25)
                    + (REAL)5.0 / (REAL)24.0) + (REAL)55.0 / (REAL)96.0;
    if (u < (REAL)2.5) { u -= (REAL)2.5; u *= u; return u * u * ((REAL)1.0 / (REAL)24.0); }
    return 0;
}

/* first derivatives; the sign factor is an int, as in the reference */
static REAL FN(dbs2)(REAL u)
{
    int sg = (u > 0 ? 1 : -1);
    u = (REAL)fabs(u);
    if (u < (REAL)0.5) return -2 * u * sg;
    if (u < (REAL)1.5) return (u - (REAL)1.5) * sg;
    return 0;
}

static REAL FN(dbs3)(REAL u)
{
    int sg = (u > 0 ? 1 : -1);
    u = (REAL)fabs(u);
    if (u < 1) return u * ((REAL)1.5 * u - (REAL)2.0) * sg;
    if (u < 2) { u = u - (REAL)2.0; return (REAL)-0.5 * u * u * sg; }
    return 0;
}

static REAL FN(dbs4)(REAL u)
{
    int sg = (u > 0 ? 1 : -1);
    u = (REAL)fabs(u);
    if (u < (REAL)0.5) return (u * (u * u - (REAL)5.0 / (REAL)4.0)) * sg;
    if (u < (REAL)1.5)
        return (u * (u * (u * ((REAL)-2.0 / (REAL)3.0) + (REAL)2.5) - (REAL)5.0 / (REAL)2.0)
                + (REAL)5.0 / (REAL)24.0) * sg;
    if (u < (REAL)2.5) { u = u * (REAL)2.0 - (REAL)5.0; return ((REAL)1.0 / (REAL)48.0) * u * u * u * sg; }
    return 0;
}

/* second derivatives */
static REAL FN(hbs3)(REAL u)
{
    u = (REAL)fabs(u);
    if (u < 1) return (REAL)3.0 * u - (REAL)2.0;
    if (u < 2) return (REAL)2.0 - u;
    return 0;
}

static REAL FN(hbs4)(REAL u)
{
    u = (REAL)fabs(u);
    if (u < (REAL)0.5) return (REAL)3.0 * u * u - (REAL)5.0 / (REAL)4.0;
    if (u < (REAL)1.5) return u * ((REAL)-2.0 * u + (REAL)5.0) - (REAL)5.0 / (REAL)2.0;
    if (u < (REAL)2.5) { u = u * (REAL)2.0 - (REAL)5.0; return u * u / (REAL)8.0; }
    return 0;
}

/* Fractional offset t = x - o for the knot o of ppo_knot(); the FP32 instance rounds
 * it in float like bspline.cu:67, the FP64 instance keeps it exact. */
static REAL FN(frac)(float x, long long o)
{
#if PPO_IS_DOUBLE
    return (double)x - (double)o;
#else
    return x - (float)o;
#endif
}

/* Weights w[k] = B(t-k), k < dp   (bspline.cu:63-91) */
static void FN(wts)(int dp, REAL t, REAL *w)
{
    int k;
    switch (dp) {
    case 1: w[0] = 1; break;
    case 3: for (k = 0; k < 3; k++) w[k] = FN(bs2)(t - (REAL)k); break;
    case 4: for (k = 0; k < 4; k++) w[k] = FN(bs3)(t - (REAL)k); break;
    case 5: for (k = 0; k < 5; k++) w[k] = FN(bs4)(t - (REAL)k); break;
    default: w[0] = (REAL)1.0 - t; w[1] = t; break; /* dp == 2 */
    }
}

/* Derivative weights (bspline.cu:143-171) */
static void FN(dwts)(int dp, REAL t, REAL *w)
{
    int k;
    switch (dp) {
    case 1: w[0] = 0; break;
    case 3: for (k = 0; k < 3; k++) w[k] = FN(dbs2)(t - (REAL)k); break;
    case 4: for (k = 0; k < 4; k++) w[k] = FN(dbs3)(t - (REAL)k); break;
    case 5: for (k = 0; k < 5; k++) w[k] = FN(dbs4)(t - (REAL)k); break;
    default: w[0] = -1; w[1] = 1; break;
    }
}

/* Second-derivative weights (bspline.cu:213-244); degree 2 uses the constant
 * [1,-2,1] stencil and degrees 0/1 give zero. */
static void FN(hwts)(int dp, REAL t, REAL *w)
{
    int k;
    switch (dp) {
    case 1: w[0] = 0; break;
    case 3: w[0] = 1; w[1] = -2; w[2] = 1; break;
    case 4: for (k = 0; k < 4; k++) w[k] = FN(hbs3)(t - (REAL)k); break;
    case 5: for (k = 0; k < 5; k++) w[k] = FN(hbs4)(t - (REAL)k); break;
    default: w[0] = 0; w[1] = 0; break;
    }
}

/* One axis of a sample: knot, weights, folded indices and signs. */
typedef struct {
    REAL w[PPO_MAXD], dw[PPO_MAXD], hw[PPO_MAXD];
    u64 b[PPO_MAXD];
    int s[PPO_MAXD];
} FN(axis_t);

static void FN(axis)(float x, int dp, u64 d, int code, int want_d, int want_h, FN(axis_t) *a)
{
    long long o = ppo_knot(x, dp);
    REAL t = FN(frac)(x, o);
    int i;
    FN(wts)(dp, t, a->w);
    if (want_d) FN(dwts)(dp, t, a->dw);
    if (want_h) FN(hwts)(dp, t, a->hw);
    for (i = 0; i < dp; i++) ppo_fold(code, (long long)i + o, d, &a->b[i], &a->s[i]);
}

/* Gather the dp0 x dp1 x dp2 patch, k outer / j / i inner, sign applied as an int
 * factor (patch.cu:62-83). */
static void FN(gather)(const int *dp, const FN(axis_t) *A, const u64 *d0, const REAL *f0, REAL *fp)
{
    int i, j, k;
    for (k = 0; k < dp[2]; k++) {
        u64 tk = d0[1] * A[2].b[k];
        int sk = A[2].s[k];
        for (j = 0; j < dp[1]; j++) {
            u64 tj = d0[0] * (A[1].b[j] + tk);
            int s = sk * A[1].s[j];
            for (i = 0; i < dp[0]; i++) *fp++ = f0[A[0].b[i] + tj] * (int)(s * A[0].s[i]);
        }
    }
}

/* Separable contraction of a patch with weights a(x) b(y) c(z) (pushpull_dev.cu:7-26) */
static REAL FN(contract)(const int *dp, const REAL *fp, const REAL *a, const REAL *b, const REAL *c)
{
    int i, j, k;
    REAL acc = 0;
    for (k = 0; k < dp[2]; k++) {
        REAL ck = c[k];
        for (j = 0; j < dp[1]; j++) {
            REAL bj = ck * b[j];
            for (i = 0; i < dp[0]; i++) acc += a[i] * bj * *fp++;
        }
    }
    return acc;
}

/* Gradient in one sweep (pushpull_dev.cu:33-58) */
static void FN(contract_grad)(const int *dp, const REAL *fp, const FN(axis_t) *A, REAL *g)
{
    int i, j, k;
    g[0] = g[1] = g[2] = 0;
    for (k = 0; k < dp[2]; k++) {
        REAL wk = A[2].w[k], dwk = A[2].dw[k];
        for (j = 0; j < dp[1]; j++) {
            REAL cx = wk * A[1].w[j], cy = wk * A[1].dw[j], cz = dwk * A[1].w[j];
            for (i = 0; i < dp[0]; i++) {
                REAL v = *fp++;
                g[0] += A[0].dw[i] * cx * v;
                v *= A[0].w[i];
                g[1] += cy * v;
                g[2] += cz * v;
            }
        }
    }
}

/* Scatter s * wa(x) wb(y) wc(z) * val into f0, same order and rounding as
 * weight_patch + put_patch (pushpull_dev.cu:119-137, patch.cu:110-139). */
static void FN(scatter)(const int *dp, const FN(axis_t) *A, const u64 *d0, REAL *f0,
                        const REAL *a, const REAL *b, const REAL *c, REAL val)
{
    int i, j, k;
    for (k = 0; k < dp[2]; k++) {
        u64 tk = d0[1] * A[2].b[k];
        int sk = A[2].s[k];
        REAL ck = c[k];
        for (j = 0; j < dp[1]; j++) {
            u64 tj = d0[0] * (A[1].b[j] + tk);
            int s = sk * A[1].s[j];
            REAL bj = ck * b[j];
            for (i = 0; i < dp[0]; i++) {
                REAL wt = a[i] * bj;
                REAL v = wt * val;
                f0[A[0].b[i] + tj] += v * (int)(s * A[0].s[i]);
            }
        }
    }
}

static int FN(sample)(const float *phi, u64 i, u64 n1, const u64 *d0, int ext, float *x)
{
    x[0] = phi[i] - 1.0f;
    x[1] = phi[i + n1] - 1.0f;
    x[2] = phi[i + 2 * n1] - 1.0f;
    return ppo_voxok(x, d0, ext);
}

static void FN(setup)(const float *x, const u64 *d0, const int *bnd, const int *dp, int want_d, int want_h, FN(axis_t) *A)
{
    int a;
    for (a = 0; a < 3; a++) FN(axis)(x[a], dp[a], d0[a], bnd[a], want_d, want_h, &A[a]);
}

/* ---- per-slice entry points, reference signatures (pushpull.c:14-152) ----- */

void FN(ppo_pull)(REAL *f1, const float *phi, const REAL *f0, const u64 *d0, u64 n1,
                  const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    REAL fp[PPO_MAXD * PPO_MAXD * PPO_MAXD];
    FN(axis_t) A[3];
    u64 i;
    float x[3];
    for (i = 0; i < n1; i++) {
        if (!FN(sample)(phi, i, n1, d0, ext, x)) { f1[i] = (REAL)NAN; continue; }
        FN(setup)(x, d0, bnd, dp, 0, 0, A);
        FN(gather)(dp, A, d0, f0, fp);
        f1[i] = FN(contract)(dp, fp, A[0].w, A[1].w, A[2].w);
    }
}

void FN(ppo_pullg)(REAL *g1, const float *phi, const REAL *f0, const u64 *d0, u64 n1,
                   const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    REAL fp[PPO_MAXD * PPO_MAXD * PPO_MAXD], g[3];
    FN(axis_t) A[3];
    u64 i;
    float x[3];
    for (i = 0; i < n1; i++) {
        if (!FN(sample)(phi, i, n1, d0, ext, x)) {
            g1[i] = g1[i + n1] = g1[i + 2 * n1] = (REAL)NAN;
            continue;
        }
        FN(setup)(x, d0, bnd, dp, 1, 0, A);
        FN(gather)(dp, A, d0, f0, fp);
        FN(contract_grad)(dp, fp, A, g);
        g1[i] = g[0];
        g1[i + n1] = g[1];
        g1[i + 2 * n1] = g[2];
    }
}

/* Hessian, column-major 3x3 with symmetric copies (pushpull.c:70-102) */
void FN(ppo_pullh)(REAL *h1, const float *phi, const REAL *f0, const u64 *d0, u64 n1,
                   const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    REAL fp[PPO_MAXD * PPO_MAXD * PPO_MAXD];
    FN(axis_t) A[3];
    u64 i;
    int c;
    float x[3];
    for (i = 0; i < n1; i++) {
        REAL hxx, hyy, hzz, hxy, hxz, hyz;
        if (!FN(sample)(phi, i, n1, d0, ext, x)) {
            for (c = 0; c < 9; c++) h1[i + c * n1] = (REAL)NAN;
            continue;
        }
        FN(setup)(x, d0, bnd, dp, 1, 1, A);
        FN(gather)(dp, A, d0, f0, fp);
        hxx = FN(contract)(dp, fp, A[0].hw, A[1].w, A[2].w);
        hyy = FN(contract)(dp, fp, A[0].w, A[1].hw, A[2].w);
        hzz = FN(contract)(dp, fp, A[0].w, A[1].w, A[2].hw);
        hxy = FN(contract)(dp, fp, A[0].dw, A[1].dw, A[2].w);
        hxz = FN(contract)(dp, fp, A[0].dw, A[1].w, A[2].dw);
        hyz = FN(contract)(dp, fp, A[0].w, A[1].dw, A[2].dw);
        h1[i] = hxx;
        h1[i + 4 * n1] = hyy;
        h1[i + 8 * n1] = hzz;
        h1[i + 1 * n1] = h1[i + 3 * n1] = hxy;
        h1[i + 2 * n1] = h1[i + 6 * n1] = hxz;
        h1[i + 5 * n1] = h1[i + 7 * n1] = hyz;
    }
}

/* Push: accumulate into f0 (caller zeroes it); skip non-finite f1 and invalid x
 * (pushpull.c:106-123). Sequential, so deterministic. */
void FN(ppo_push)(REAL *f0, const float *phi, const REAL *f1, const u64 *d0, u64 n1,
                  const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    FN(axis_t) A[3];
    u64 i;
    float x[3];
    for (i = 0; i < n1; i++) {
        REAL v = f1[i];
        if (!isfinite(v)) continue;
        if (!FN(sample)(phi, i, n1, d0, ext, x)) continue;
        FN(setup)(x, d0, bnd, dp, 0, 0, A);
        FN(scatter)(dp, A, d0, f0, A[0].w, A[1].w, A[2].w, v);
    }
}

/* Push of gradients: the adjoint of ppo_pullg. Follows pushpull.cu:145-163 (skip
 * only invalid x; three scatter passes). */
void FN(ppo_pushg)(REAL *f0, const float *phi, const REAL *g1, const u64 *d0, u64 n1,
                   const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    FN(axis_t) A[3];
    u64 i;
    float x[3];
    for (i = 0; i < n1; i++) {
        if (!FN(sample)(phi, i, n1, d0, ext, x)) continue;
        FN(setup)(x, d0, bnd, dp, 1, 0, A);
        FN(scatter)(dp, A, d0, f0, A[0].dw, A[1].w, A[2].w, g1[i]);
        FN(scatter)(dp, A, d0, f0, A[0].w, A[1].dw, A[2].w, g1[i + n1]);
        FN(scatter)(dp, A, d0, f0, A[0].w, A[1].w, A[2].dw, g1[i + 2 * n1]);
    }
}

/* Affine variants (pushpull.c:176-244): coordinates x = A*(i,j,k,1) over the output
 * grid d1 from the 0-based 3x4 matrix A (column-major), in the reference's order
 * xj = (A3 j + A6 k) + A9, x = xj + A0 i, all in FP32. */
static void FN(affine_coords)(const float *A, int i, int j, int k, float *x)
{
    float xj0 = A[3] * j + A[6] * k + A[9];
    float xj1 = A[4] * j + A[7] * k + A[10];
    float xj2 = A[5] * j + A[8] * k + A[11];
    x[0] = xj0 + A[0] * i;
    x[1] = xj1 + A[1] * i;
    x[2] = xj2 + A[2] * i;
}

void FN(ppo_pull_affine)(REAL *f1, const REAL *f0, const u64 *d0, u64 n1, const u64 *d1, const float *A,
                         const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    REAL fp[PPO_MAXD * PPO_MAXD * PPO_MAXD];
    FN(axis_t) Ax[3];
    int i, j, k;
    u64 v = 0;
    float x[3];
    (void)n1;
    for (k = 0; k < (int)d1[2]; k++)
        for (j = 0; j < (int)d1[1]; j++)
            for (i = 0; i < (int)d1[0]; i++, v++) {
                FN(affine_coords)(A, i, j, k, x);
                if (!ppo_voxok(x, d0, ext)) { f1[v] = (REAL)NAN; continue; }
                FN(setup)(x, d0, bnd, dp, 0, 0, Ax);
                FN(gather)(dp, Ax, d0, f0, fp);
                f1[v] = FN(contract)(dp, fp, Ax[0].w, Ax[1].w, Ax[2].w);
            }
}

void FN(ppo_push_affine)(REAL *f0, const REAL *f1, const u64 *d0, u64 n1, const u64 *d1, const float *A,
                         const int *bnd, const u64 *dpu, int ext)
{
    int dp[3] = {(int)dpu[0], (int)dpu[1], (int)dpu[2]};
    FN(axis_t) Ax[3];
    int i, j, k;
    u64 v = 0;
    float x[3];
    (void)n1;
    for (k = 0; k < (int)d1[2]; k++)
        for (j = 0; j < (int)d1[1]; j++)
            for (i = 0; i < (int)d1[0]; i++, v++) {
                REAL val = f1[v];
                if (!isfinite(val)) continue;
                FN(affine_coords)(A, i, j, k, x);
                if (!ppo_voxok(x, d0, ext)) continue;
                FN(setup)(x, d0, bnd, dp, 0, 0, Ax);
                FN(scatter)(dp, Ax, d0, f0, Ax[0].w, Ax[1].w, Ax[2].w, val);
            }
}

#undef FN
