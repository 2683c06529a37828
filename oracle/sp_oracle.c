/* sp_oracle.c -- CPU oracle of the sparse regulariser (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, in the same per-operation order, vel2mom and relax of
 * artifacts/C/sparse_operator.c:45-135 with their voxel routines in
 * artifacts/C/sparse_operator_dev.cu:12-204 (sp_conv, Hv, make_A, relax1/relax_padded1,
 * vel2mom1/vel2mom_padded1) and the Cholesky pair of artifacts/C/chol.cu. Only the tests
 * and the CPU-baseline leg of the benchmark call it; the product path never does.
 *
 * Sparse kernel layout (src/operator.jl:163-206, `sparsify`): for the channel pair (i, j)
 * (p = i + nchan*j) the list of entries starts at offset[p] (an index into values and
 * patch) and holds length[p] entries, the first being the centre of the stencil; the
 * entry just before the list, values[offset[p]-1], is the sum of the non-centre values.
 * patch[t] numbers a position of the dp[0] x dp[1] x dp[2] stencil box, x fastest.
 *
 * Every stencil read folds the position through the input channel's boundary triple
 * (the same folds as pull / push), so the interior and the edge voxels take one path:
 * in the interior the folds are identities and the values are those the reference reads
 * through its precomputed image offsets, in the same order.
 */
#include <math.h>
#include <stddef.h>

typedef unsigned long long u64;
typedef long long i64;

void ppo_fold_export(int code, long long i, u64 d, u64 *b, int *s);

typedef struct {
    int dp[3];           /* stencil box */
    u64 d[3];            /* volume */
    int nchan;           /* channels */
    int hcode;           /* relax: 0 no Hessian, 1 diagonal, 2 full symmetric */
    const int *offset;   /* nchan x nchan, p = i + nchan*j */
    const int *length;
    const float *values;
    const int *patch;
    const int *bnd;      /* 3 x nchan: channel c's triple at bnd + 3c */
} spk;

/* value of box position t of input channel c around voxel (x, y, z), folded and signed */
static float box_value(const spk *k, const float *vc, const int *bc, i64 x, i64 y, i64 z, int t)
{
    const int a = t % k->dp[0], b = (t / k->dp[0]) % k->dp[1], c = t / (k->dp[0] * k->dp[1]);
    u64 ix, iy, iz;
    int sx, sy, sz;
    ppo_fold_export(bc[0], x - k->dp[0] / 2 + a, k->d[0], &ix, &sx);
    ppo_fold_export(bc[1], y - k->dp[1] / 2 + b, k->d[1], &iy, &sy);
    ppo_fold_export(bc[2], z - k->dp[2] / 2 + c, k->d[2], &iz, &sz);
    return vc[ix + k->d[0] * (iy + k->d[1] * iz)] * (float)(sx * sy * sz);
}

/* sparse_operator_dev.cu:12-20: the list's convolution, in the differences-to-centre form
 * (the centre value times (centre + sum of the others) added last) */
static float list_conv(const spk *k, int p, const float *vc, const int *bc, i64 x, i64 y, i64 z)
{
    const float *val = k->values + k->offset[p];
    const int *pt = k->patch + k->offset[p];
    const int n = k->length[p];
    const float f0 = box_value(k, vc, bc, x, y, z, pt[0]);
    float s = 0.0f;
    int t;
    for (t = 1; t < n; t++) s += val[t] * (box_value(k, vc, bc, x, y, z, pt[t]) - f0);
    s += (val[0] + val[-1]) * f0;
    return s;
}

static i64 imin(i64 a, i64 b) { return a < b ? a : b; }
static i64 imax(i64 a, i64 b) { return a > b ? a : b; }

/* interior block [s, e) of sparse_operator.c:48-53 (voxels whose whole box lies inside) */
static void interior(const spk *k, i64 s[3], i64 e[3])
{
    int q;
    for (q = 0; q < 3; q++) {
        s[q] = imin((k->dp[q] - 1) / 2, (i64)k->d[q]);
        e[q] = imax(s[q], (i64)k->d[q] - (k->dp[q] - 1) / 2);
    }
}

/* u = L v over the volume (sparse_operator_dev.cu:164-204: output channel i sums the
 * lists (i, 0), (i, 1), ... in that order). The interior block starts each sum from zero
 * and stores it; outside it the reference adds each list's result to u in turn
 * (sparse_operator_dev.cu:200), so with accum_edges the sum starts from the old u there. */
void spo_vel2mom(float *u, const float *v, const spk *k, int accum_edges)
{
    const u64 nd = k->d[0] * k->d[1] * k->d[2];
    i64 x, y, z, s[3], e[3];
    int i, j;
    interior(k, s, e);
    for (z = 0; z < (i64)k->d[2]; z++)
        for (y = 0; y < (i64)k->d[1]; y++)
            for (x = 0; x < (i64)k->d[0]; x++) {
                const u64 m = (u64)x + k->d[0] * ((u64)y + k->d[1] * (u64)z);
                const int edge = x < s[0] || x >= e[0] || y < s[1] || y >= e[1] || z < s[2] || z >= e[2];
                for (i = 0; i < k->nchan; i++) {
                    float t = (edge && accum_edges) ? u[m + nd * i] : 0.0f;
                    for (j = 0; j < k->nchan; j++)
                        t += list_conv(k, i + k->nchan * j, v + nd * j, k->bnd + 3 * j, x, y, z);
                    u[m + nd * i] = t;
                }
            }
}

/* chol.cu: Cholesky of the n x n matrix a (row-major triangle, diagonal in p), then the
 * two triangular solves */
static void chol_factor(int n, float *a, float *p)
{
    int i, j, q;
    for (i = 0; i < n; i++)
        for (j = i; j < n; j++) {
            float t = a[i * n + j];
            for (q = i - 1; q >= 0; q--) t -= a[i * n + q] * a[j * n + q];
            if (j == i) p[i] = sqrtf(t);
            else a[j * n + i] = t / p[i];
        }
}

static void chol_solve(int n, const float *a, const float *p, const float *b, float *x)
{
    int i, q;
    for (i = 0; i < n; i++) {
        float t = b[i];
        for (q = i - 1; q >= 0; q--) t -= a[i * n + q] * x[q];
        x[i] = t / p[i];
    }
    for (i = n - 1; i >= 0; i--) {
        float t = x[i];
        for (q = i + 1; q < n; q++) t -= a[q * n + i] * x[q];
        x[i] = t / p[i];
    }
}

/* one Gauss-Seidel update of voxel (x, y, z): v += (L_centre + H)^-1 (g - L v - H v)
 * (sparse_operator_dev.cu:89-161) */
static void relax_voxel(float *v, const float *g, const float *h, const spk *k, i64 x, i64 y, i64 z)
{
    const u64 nd = k->d[0] * k->d[1] * k->d[2];
    const u64 m = (u64)x + k->d[0] * ((u64)y + k->d[1] * (u64)z);
    const int n = k->nchan;
    float b[8], A[64], p[8], sol[8];
    int i, j, ii;
    for (i = 0; i < n; i++) b[i] = g[m + nd * i];
    for (j = 0; j < n; j++)
        for (i = 0; i < n; i++) b[i] -= list_conv(k, i + n * j, v + nd * j, k->bnd + 3 * j, x, y, z);
    if (k->hcode) {  /* Hv */
        for (ii = 0; ii < n; ii++) b[ii] -= h[m + nd * ii] * v[m + nd * ii];
        if (k->hcode == 2)
            for (j = 0, ii = n; j < n; j++)
                for (i = j + 1; i < n; i++, ii++) {
                    b[j] -= h[m + nd * ii] * v[m + nd * i];
                    b[i] -= h[m + nd * ii] * v[m + nd * j];
                }
    }
    for (j = 0; j < n; j++) {  /* make_A: the stencils' centres, plus H */
        A[j * n + j] = k->values[k->offset[j * n + j]] * 1.000001f;
        for (i = j + 1; i < n; i++) A[i + j * n] = A[i * n + j] = k->values[k->offset[i + j * n]];
    }
    if (k->hcode) {
        for (ii = 0; ii < n; ii++) A[ii + n * ii] += h[m + nd * ii] * 1.000001f;
        if (k->hcode == 2)
            for (j = 0, ii = n; j < n; j++)
                for (i = j + 1; i < n; i++, ii++) A[i + j * n] = A[j + i * n] += h[m + nd * ii];
    }
    chol_factor(n, A, p);
    chol_solve(n, A, p, b, sol);
    for (i = 0; i < n; i++) v[m + nd * i] += sol[i];
}

/* the voxels of [lo, hi) colour by colour: colour (a, b, c) holds the voxels whose offsets
 * from lo are multiples of dp along each axis (sparse_operator.c:66-114) */
static void relax_region(float *v, const float *g, const float *h, const spk *k, const i64 lo[3], const i64 hi[3])
{
    i64 a, b, c, x, y, z;
    for (c = lo[2]; c < imin(hi[2], lo[2] + k->dp[2]); c++)
        for (b = lo[1]; b < imin(hi[1], lo[1] + k->dp[1]); b++)
            for (a = lo[0]; a < imin(hi[0], lo[0] + k->dp[0]); a++)
                for (z = c; z < hi[2]; z += k->dp[2])
                    for (y = b; y < hi[1]; y += k->dp[1])
                        for (x = a; x < hi[0]; x += k->dp[0]) relax_voxel(v, g, h, k, x, y, z);
}

/* nit sweeps; each: the six edge slabs, then the middle block (sparse_operator.c:116-135) */
void spo_relax(float *v, const float *g, const float *h, const spk *k, int nit)
{
    i64 s[3], e[3], d[3];
    int q, it;
    interior(k, s, e);
    for (q = 0; q < 3; q++) d[q] = (i64)k->d[q];
    for (it = 0; it < nit; it++) {
        const i64 r[7][6] = {
            {0, d[0], 0, d[1], 0, s[2]},    {0, d[0], 0, s[1], s[2], e[2]},    {0, s[0], s[1], e[1], s[2], e[2]},
            {e[0], d[0], s[1], e[1], s[2], e[2]}, {0, d[0], e[1], d[1], s[2], e[2]}, {0, d[0], 0, d[1], e[2], d[2]},
            {s[0], e[0], s[1], e[1], s[2], e[2]}};
        for (q = 0; q < 7; q++) {
            const i64 lo[3] = {r[q][0], r[q][2], r[q][4]}, hi[3] = {r[q][1], r[q][3], r[q][5]};
            relax_region(v, g, h, k, lo, hi);
        }
    }
}

/* flat entry points for ctypes: dims = (d0, d1, d2), dp = stencil box */
void spo_vel2mom_export(float *u, const float *v, const u64 *dims, int nchan, const int *offset, const int *length,
                        const float *values, const int *patch, const int *dp, const int *bnd, int accum_edges)
{
    spk k = {{dp[0], dp[1], dp[2]}, {dims[0], dims[1], dims[2]}, nchan, 0, offset, length, values, patch, bnd};
    spo_vel2mom(u, v, &k, accum_edges);
}

void spo_relax_export(float *v, const float *g, const float *h, int hcode, const u64 *dims, int nchan,
                      const int *offset, const int *length, const float *values, const int *patch, const int *dp,
                      const int *bnd, int nit)
{
    spk k = {{dp[0], dp[1], dp[2]}, {dims[0], dims[1], dims[2]}, nchan, hcode, offset, length, values, patch, bnd};
    spo_relax(v, g, h, &k, nit);
}
