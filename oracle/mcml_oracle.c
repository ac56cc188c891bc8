/*
 * mcml_oracle.c -- CPU oracle (test infrastructure only; see mcml_oracle.h).
 *
 * A statement-by-statement restatement of the active path of the reference
 * kernel src/Simulation/MCLightPropagation.cl and of the host binning in
 * src/Simulation/CalculatePropagation.cpp.  Built with -ffp-contract=off so
 * that every float operation is one IEEE-754 operation in source order;
 * division and square root are the correctly rounded C operators.
 *
 * Where the reference relies on OpenCL built-ins whose results are not fixed
 * by the language (log, cos, acos, normalize, dot), this file fixes them:
 *   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   normalize(p) = p * (1/sqrt(dot(p,p))), and p itself when dot(p,p) == 0
 *   sign(x)      = OpenCL sign(): +-1, +-0 for +-0, 0 for NaN
 *   pown(x,2)    = x*x
 *   log, cos, acos: orc_logf / orc_cosf / orc_acosf below.
 *
 * Reference semantics that are reproduced (they define its results): the
 * signed-cosine Fresnel, the kill-on-(mua==0 || mus==0) exit, the distMin
 * R/T split, the (1-Rsp) factor at exit on top of w0 = 1-Rsp, the near-axis
 * normal override, refTri exclusion right after a boundary hit, lowest-index
 * tie-break, VScatter's rotation construction, the xi order, int() binning.
 * Reference defects that are NOT reproduced: the racy, unchecked Arz
 * read-modify-write and its cumulative host readback, the shared lock-ordered
 * RNG, and the mediumIDs[-1] read when no face lies ahead (here: "escaped").
 */
#include "mcml_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* float helpers                                                        */
/* ------------------------------------------------------------------ */
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

typedef struct { float x, y, z; } of3;

static inline of3 V(float x, float y, float z) { of3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline of3 vadd(of3 a, of3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline of3 vsub(of3 a, of3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline of3 vmul(float s, of3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline of3 vneg(of3 a) { return V(-a.x, -a.y, -a.z); }
static inline float vdot(of3 a, of3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline of3 vcross(of3 a, of3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline of3 vnormalize(of3 p)
{
    float d = vdot(p, p);
    if (d == 0.0f) return p;
    float inv = 1.0f / sqrtf(d);
    return vmul(inv, p);
}
static inline of3 vload3(const float *a, int64_t i) { return V(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

/* 1 - c^2 is negative only when |c| exceeds 1 by rounding (two unit
 * vectors whose squared norms round above 1).  The reference takes the
 * square root anyway (MCLightPropagation.cl:520, 541, 880) and the photon
 * continues as NaN; here the argument is clamped at 0, which leaves every
 * finite trajectory unchanged. */
static inline float orc_clamp0(float q) { return q < 0.0f ? 0.0f : q; }

float orc_sign(float x)
{
    if (x > 0.0f) return 1.0f;
    if (x < 0.0f) return -1.0f;
    if (x == 0.0f) return x; /* keeps the sign of zero */
    return 0.0f;             /* NaN */
}

/* ------------------------------------------------------------------ */
/* Transcendentals                                                      */
/* ------------------------------------------------------------------ */
/* Natural logarithm.  Cody & Waite argument reduction x = 2^e * m with
 * m in [sqrt(2)/2, sqrt(2)], then log(m) = 2 atanh(s), s = (m-1)/(m+1),
 * |s| <= 0.1716, from the Taylor series of atanh truncated after s^9
 * (Abramowitz & Stegun 4.1.27; truncation error < 2.2e-9 relative).
 * ln 2 = LN2_HI + LN2_LO with LN2_HI holding 12 significant bits, so
 * e * LN2_HI is exact. */
#define ORC_LN2_HI 0.693115234375f
#define ORC_LN2_LO 3.19461833e-05f
float orc_logf(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    uint32_t ix = f2u(x);
    if (ix < 0x00800000u) { /* subnormal: scale into the normal range */
        x = x * 8388608.0f; /* 2^23, exact */
        ix = f2u(x);
        e = -23;
    }
    e += (int)(ix >> 23) - 127;
    float m = u2f((ix & 0x007fffffu) | 0x3f800000u); /* [1, 2) */
    if (m > 1.41421354f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = z * (0.333333343f + z * (0.200000003f + z * (0.142857149f + z * 0.111111112f)));
    float logm = 2.0f * s + (2.0f * s) * p;
    float fe = (float)e;
    return fe * ORC_LN2_HI + (fe * ORC_LN2_LO + logm);
}

/* Cosine for |x| < 2^10.  Cody & Waite reduction by pi/2 with a three-part
 * pi/2 (PIO2_1, PIO2_2 hold 12 significant bits each, so k*PIO2_i is exact
 * for k < 2^12), then the Taylor series of cos (through r^10) and sin
 * (through r^9) on |r| <= pi/4, selected by quadrant.  Truncation error
 * < 1.2e-10 (cos) and < 2e-9 relative (sin). */
#define ORC_PIO2_1 1.5703125f
#define ORC_PIO2_2 0.000483751297f
#define ORC_PIO2_3 7.54978995e-08f
#define ORC_2_OVER_PI 0.636619747f
static inline float orc_cos_poly(float r)
{
    float z = r * r;
    return 1.0f + z * (-0.5f + z * (0.0416666679f + z * (-0.00138888892f +
                 z * (2.48015876e-05f + z * (-2.75573188e-07f)))));
}
static inline float orc_sin_poly(float r)
{
    float z = r * r;
    return r + (r * z) * (-0.166666672f + z * (0.00833333377f + z * (-0.000198412701f +
                 z * 2.75573188e-06f)));
}
float orc_cosf(float x)
{
    if (x != x) return x;
    float ax = fabsf(x);
    if (!(ax < 1024.0f)) return u2f(0x7fc00000u); /* outside the supported domain */
    int k = (int)(ax * ORC_2_OVER_PI + 0.5f);
    float fk = (float)k;
    float r = ((ax - fk * ORC_PIO2_1) - fk * ORC_PIO2_2) - fk * ORC_PIO2_3;
    switch (k & 3) {
    case 0: return orc_cos_poly(r);
    case 1: return -orc_sin_poly(r);
    case 2: return -orc_cos_poly(r);
    default: return orc_sin_poly(r);
    }
}

/* Arc cosine.  asin on |x| <= 1/2 from its Maclaurin series
 * asin x = sum (2n)! / (4^n (n!)^2 (2n+1)) x^(2n+1), n <= 10 (A&S 4.4.40;
 * tail < 2e-9 relative), and the half-angle identities
 * acos x = 2 asin(sqrt((1-x)/2)) for x > 1/2,
 * acos x = pi - 2 asin(sqrt((1+x)/2)) for x < -1/2,
 * acos x = pi/2 - asin x otherwise. */
#define ORC_PIO2_HI 1.57079637f
#define ORC_PIO2_LO (-4.37113883e-08f)
#define ORC_PI_HI 3.14159274f
#define ORC_PI_LO (-8.74227766e-08f)
static inline float orc_asin_poly(float x)
{
    float z = x * x;
    float p = 0.166666672f + z * (0.075000003f + z * (0.0446428582f + z * (0.0303819440f +
              z * (0.0223721601f + z * (0.0173527617f + z * (0.0139648430f + z * (0.0115518477f +
              z * (0.00976160448f + z * 0.00839033350f))))))));
    return x + (x * z) * p;
}
float orc_acosf(float x)
{
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return u2f(0x7fc00000u);
    if (x > 0.5f) {
        float s = sqrtf((1.0f - x) * 0.5f);
        return 2.0f * orc_asin_poly(s);
    }
    if (x < -0.5f) {
        float s = sqrtf((1.0f + x) * 0.5f);
        return ORC_PI_HI - (2.0f * orc_asin_poly(s) - ORC_PI_LO);
    }
    return ORC_PIO2_HI - (orc_asin_poly(x) - ORC_PIO2_LO);
}

/* ------------------------------------------------------------------ */
/* Random numbers                                                       */
/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random
 * numbers: as easy as 1, 2, 3"). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

#define ORC_MBIG 1000000000
#define ORC_MSEED 161803398
#define ORC_MZ 0

/* PSNR::InitRand3 (PSNR.cpp:48-80), including its first draw. */
void orc_ran3_init(int64_t idum, int64_t ma[56], int32_t *inext, int32_t *inextp)
{
    int64_t mj, mk;
    int i, ii, k;
    for (i = 0; i < 56; ++i) ma[i] = 0;
    mj = ORC_MSEED - (idum < 0 ? -idum : idum);
    mj %= ORC_MBIG;
    ma[55] = mj;
    mk = 1;
    for (i = 1; i <= 54; i++) {
        ii = (21 * i) % 55;
        ma[ii] = mk;
        mk = mj - mk;
        if (mk < ORC_MZ) mk += ORC_MBIG;
        mj = ma[ii];
    }
    for (k = 1; k <= 4; k++)
        for (i = 1; i <= 55; i++) {
            ma[i] -= ma[1 + (i + 30) % 55];
            if (ma[i] < ORC_MZ) ma[i] += ORC_MBIG;
        }
    *inext = 0;
    *inextp = 31;
    (void)orc_ran3_next(ma, inext, inextp);
}

/* PSNR::GetRan3 (PSNR.cpp:105-117) integer step. */
int64_t orc_ran3_next(int64_t ma[56], int32_t *inext, int32_t *inextp)
{
    if (++(*inext) == 56) *inext = 1;
    if (++(*inextp) == 56) *inextp = 1;
    int64_t mj = ma[*inext] - ma[*inextp];
    if (mj < ORC_MZ) mj += ORC_MBIG;
    ma[*inext] = mj;
    return mj;
}

/* DoRand3 (.cl:1204-1243): the float-state continuation used on the
 * device.  MB = 1e9 is exact in float. */
float orc_ran3f_next(float s[56], int32_t *inext, int32_t *inextp)
{
    int32_t i = *inext + 1, j = *inextp + 1;
    if (i == 56) i = 1;
    if (j == 56) j = 1;
    float r = s[i] - s[j];
    if (r < 0.0f) r = r + 1000000000.0f;
    s[i] = r;
    *inext = i;
    *inextp = j;
    return r / 1000000000.0f;
}

typedef struct {
    int mode;
    uint32_t key[2];
    uint32_t id[2];
    uint32_t block;
    uint32_t buf[4];
    int used;
    float s[56];
    int32_t inext, inextp;
    int64_t ndraw;
} orc_rng;

static void rng_init(orc_rng *g, int mode, uint64_t seed, int64_t photon)
{
    g->mode = mode;
    g->ndraw = 0;
    if (mode == ORC_RNG_RAN3F) {
        int64_t ma[56];
        uint64_t sd = (seed + (uint64_t)photon) % (uint64_t)ORC_MSEED;
        orc_ran3_init(-(int64_t)sd, ma, &g->inext, &g->inextp);
        for (int i = 0; i < 56; ++i) g->s[i] = (float)ma[i];
    } else {
        g->key[0] = (uint32_t)seed;
        g->key[1] = (uint32_t)(seed >> 32);
        g->id[0] = (uint32_t)photon;
        g->id[1] = (uint32_t)((uint64_t)photon >> 32);
        g->block = 0;
        g->used = 4;
    }
}

/* Uniform in [0, 1): Philox word >> 8 times 2^-24 (exact), or ran3f. */
static float rng_next(orc_rng *g)
{
    g->ndraw++;
    if (g->mode == ORC_RNG_RAN3F) return orc_ran3f_next(g->s, &g->inext, &g->inextp);
    if (g->used == 4) {
        uint32_t ctr[4] = {g->block, g->id[0], g->id[1], 0u};
        orc_philox4x32_10(ctr, g->key, g->buf);
        g->block++;
        g->used = 0;
    }
    uint32_t u = g->buf[g->used++];
    return (float)(u >> 8) * 5.96046448e-08f;
}

/* ------------------------------------------------------------------ */
/* Geometry                                                             */
/* ------------------------------------------------------------------ */
/* FindTriangleInPath (.cl:1037-1093). */
int32_t orc_find_triangle(const float *tri, int32_t ntri, const float pp[3],
                          const float vv[3], int32_t ref_tri, float *t_out)
{
    of3 p = V(pp[0], pp[1], pp[2]), v = V(vv[0], vv[1], vv[2]);
    float best = 10000.0f;
    int32_t id = -1;
    for (int32_t j = 0; j < ntri; ++j) {
        if (j == ref_tri) continue;
        of3 K = vload3(tri, 3 * (int64_t)j);
        of3 L = vload3(tri, 3 * (int64_t)j + 1);
        of3 M = vload3(tri, 3 * (int64_t)j + 2);
        of3 E = vsub(K, M), F = vsub(L, M), G = vsub(p, M);
        of3 GxE = vcross(G, E), vxF = vcross(v, F);
        float tx = vdot(GxE, F);
        float w = vdot(vxF, E);
        float ty = vdot(vxF, G);
        float tz = vdot(GxE, v);
        float t = tx / w, k = ty / w, l = tz / w;
        if (t < best && t > 0.0f && k >= 0.0f && l >= 0.0f && (k + l) < 1.0f) {
            best = t;
            id = j;
        }
    }
    *t_out = best;
    return id;
}

/* Rspecular (.cl:273-299). */
float orc_rspecular(const float *nm, int32_t ntissue)
{
    float n1 = nm[0];
    float n2 = nm[4], mua1 = nm[5], mus1 = nm[6];
    float temp = (n1 - n2) / (n1 + n2);
    float r1 = temp * temp;
    if (mua1 == 0.0f && mus1 == 0.0f && ntissue > 2) {
        float n3 = nm[8];
        temp = (n2 - n3) / (n2 + n3);
        float r2 = temp * temp;
        r1 = r1 + (1 - r1) * (1 - r1) * r2 / (1 - r1 * r2);
    }
    return r1;
}

/* GetFresnel (.cl:857-907), cos_in signed as the caller passes it. */
float orc_fresnel(float n1, float n2, float cos_in)
{
    float r;
    if (n1 == n2) {
        r = 0.0f;
    } else if (fabsf(cos_in) >= 1.0f) { /* normal incidence (|cos| > 1 only by rounding) */
        r = (n1 - n2) / (n1 + n2);
        r *= r;
    } else if (cos_in == 0.0f) {
        r = 1.0f;
    } else {
        float sin_in = sqrtf(orc_clamp0(1 - cos_in * cos_in));
        float sin_out = n1 * sin_in / n2;
        if (sin_out >= 1.0f) {
            r = 1.0f;
        } else {
            float cos_out = sqrtf(1 - sin_out * sin_out);
            float cap = cos_in * cos_out - sin_in * sin_out;
            float cam = cos_in * cos_out + sin_in * sin_out;
            float sap = sin_in * cos_out + cos_in * sin_out;
            float sam = sin_in * cos_out - cos_in * sin_out;
            r = 0.5f * (sam * sam) * ((cam * cam) + (cap * cap)) / ((sap * sap) * (cam * cam));
        }
    }
    return r;
}

/* Rotate (.cl:697-708): Rodrigues rotation of vec about axis. */
static inline of3 orc_rotate(of3 vec, of3 axis, float c, float s)
{
    float nDx = vdot(axis, vec);
    of3 nXx = vcross(axis, vec);
    of3 nXxXn = vcross(nXx, axis);
    return vadd(vadd(vmul(nDx, axis), vmul(c, nXxXn)), vmul(s, nXx));
}

static of3 vscatter(of3 v, float g, float xt, float xp)
{
    of3 e_rot = V(0.0f, 0.0f, 0.0f);
    if (v.z != 0.0f) {
        float z = -(v.x + v.y) / v.z;
        e_rot = V(1.0f, 1.0f, z);
    } else if (v.y != 0.0f) {
        float y = -(v.x) / v.y;
        e_rot = V(1.0f, y, 0.0f);
    } else if (v.x != 0.0f) {
        e_rot = V(0.0f, 1.0f, 0.0f);
    }
    e_rot = vnormalize(e_rot);
    float cos_t;
    float g1 = (1 - g * g) / (1 - g + 2 * g * xt);
    if (g != 0.0f) {
        cos_t = (1 + g * g - g1 * g1) / (2 * g);
        if (cos_t < -1) cos_t = -1;
        else if (cos_t > 1) cos_t = 1;
    } else {
        cos_t = 2 * xt - 1;
    }
    float sin_t = sqrtf(1.0f - cos_t * cos_t);
    float psi = 6.28318548f * xp; /* 2*M_PI_F, exact doubling */
    float cos_p = orc_cosf(psi);
    float sin_p = sqrtf(1.0f - cos_p * cos_p);
    if (psi > 3.14159274f) sin_p *= -1;
    of3 rt = vnormalize(orc_rotate(v, e_rot, cos_t, sin_t));
    return vnormalize(orc_rotate(rt, v, cos_p, sin_p));
}

/* VScatter (.cl:597-655). */
void orc_vscatter(float v[3], float g, float xt, float xp)
{
    of3 r = vscatter(V(v[0], v[1], v[2]), g, xt, xp);
    v[0] = r.x; v[1] = r.y; v[2] = r.z;
}

/* ------------------------------------------------------------------ */
/* Layout and binning                                                   */
/* ------------------------------------------------------------------ */
/* x86 cvtt semantics of the reference host's int(): truncation toward
 * zero, and INT_MIN for NaN or out of range. */
static inline int32_t orc_cvt(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

#define FX32 4294967296.0f
#define FX24 16777216.0f
static inline int64_t fx32(float w) { return (int64_t)(w * FX32); }
static inline int64_t fx24(float w) { return (int64_t)(w * FX24); }

void orc_layout_compute(const orc_params *p, orc_layout *L)
{
    memset(L, 0, sizeof(*L));
    /* InitLogVectors (CalculatePropagation.cpp:127-131): float division,
     * int() truncation; dimR from integer halves in double. */
    L->dimx = orc_cvt(p->model_dims[0] / p->det_grid[0]);
    L->dimy = orc_cvt(p->model_dims[1] / p->det_grid[1]);
    L->dimz = orc_cvt(p->model_dims[2] / p->det_grid[2]);
    L->dimAlpha = orc_cvt(90 / p->det_grid[5]);
    L->dimR = (int32_t)ceil(sqrt(pow((double)(L->dimx / 2), 2) + pow((double)(L->dimy / 2), 2)));
    L->sft_nx = 2 * (L->dimx / 2) + 1;
    L->sft_ny = 2 * (L->dimy / 2) + 1;
    L->arz_nr = 2 * L->dimR + 1;
    L->arz_nz = L->dimz + 1;
    /* distMin (.cl:1351) with DimsXYZ[2] = float(dimz) (cpp:236) */
    L->dist_min = p->det_grid[2] * (float)L->dimz / 2.0f;
    L->rsp = orc_rspecular(p->nmuasg, p->ntissue);
    L->n_sft = (int64_t)L->sft_nx * L->sft_ny * 3;
    L->n_rtd = (L->dimR > 0 && L->dimAlpha > 0) ? (int64_t)L->dimR * L->dimAlpha * 3 : 0;
    L->n_arz = (int64_t)L->arz_nr * L->arz_nz;
    L->n_region = p->ntissue;
    L->n_vox = (p->vox_n[0] > 0 && p->vox_n[1] > 0 && p->vox_n[2] > 0)
                   ? (int64_t)p->vox_n[0] * p->vox_n[1] * p->vox_n[2] : 0;
    L->n_rad = p->rad_n > 0 ? (int64_t)p->rad_n * 3 : 0;
    L->n_det = (int64_t)p->ndet * 2;
}

static inline void add64(int64_t *a, int64_t v)
{
    if (v == 0) return;
#pragma omp atomic
    *a += v;
}

/* CalcCoords (.cl:101-115) with CalcCartesianCoordPerAxis (.cl:34-42). */
static inline of3 calc_coords(of3 pos, of3 poi, of3 ex, of3 ey, of3 ez)
{
    of3 d = vsub(pos, poi);
    return V(vdot(vnormalize(ex), d), vdot(vnormalize(ey), d), vdot(vnormalize(ez), d));
}

/* Absorption tallies at an interaction (replaces RecordAbsorbtionRZ,
 * .cl:254-271, by a bounds-checked atomic tally). */
static void record_absorption(const orc_params *P, const orc_layout *L, orc_tallies *T,
                              int32_t tissue, float dw, float mua, of3 c, int64_t *loc)
{
    int64_t q = fx32(dw);
    loc[0] += q; /* region, flushed by the caller */
    if (T->arz) {
        float rho = sqrtf(c.x * c.x + c.y * c.y) * orc_sign(c.x) / P->det_grid[3];
        int32_t ir = orc_cvt(rho);
        int32_t iz = orc_cvt(c.z / P->det_grid[2]);
        if (ir >= -L->dimR && ir <= L->dimR && iz >= 0 && iz <= L->dimz) {
            int64_t cell = (int64_t)iz * L->arz_nr + (ir + L->dimR);
            add64(&T->arz[cell], q);
            add64(&T->phirz[cell], fx24(dw / (mua == 0.0f ? 1.0f : mua)));
        } else {
            loc[1] += q;
        }
    }
    if (T->vox && L->n_vox) {
        /* voxel index = trunc((c - origin) * (1/size)), 1/size in float */
        float fx = (c.x - P->vox_origin[0]) * (1.0f / P->vox_size[0]);
        float fy = (c.y - P->vox_origin[1]) * (1.0f / P->vox_size[1]);
        float fz = (c.z - P->vox_origin[2]) * (1.0f / P->vox_size[2]);
        int ok = fx >= 0.0f && fy >= 0.0f && fz >= 0.0f;
        int32_t ix = orc_cvt(fx), iy = orc_cvt(fy), iz = orc_cvt(fz);
        ok = ok && ix >= 0 && iy >= 0 && iz >= 0 && ix < P->vox_n[0] && iy < P->vox_n[1] && iz < P->vox_n[2];
        if (ok)
            add64(&T->vox[((int64_t)iz * P->vox_n[1] + iy) * P->vox_n[0] + ix], q);
        else
            loc[2] += q;
    }
    (void)tissue;
}

/* Host binning of one exit record (CalculatePropagation.cpp:1553-1630) and
 * the radial / detector tallies. */
static void record_exit(const orc_params *P, const orc_layout *L, orc_tallies *T, const float r8[8])
{
    if (r8[3] == 0.0f && r8[4] == 0.0f && r8[5] == 0.0f) return; /* adds zero */
    int64_t q0 = fx32(r8[3]), q1 = fx32(r8[4]), q2 = fx32(r8[5]);
    if (T->sft) {
        int32_t ex0 = -(L->dimx / 2), ex1 = L->dimx / 2, ey0 = -(L->dimy / 2), ey1 = L->dimy / 2;
        int32_t x = orc_cvt(r8[0] / P->det_grid[0]);
        int32_t y = orc_cvt(r8[1] / P->det_grid[1]);
        y = (y < ey0) ? ey0 : y;
        y = (y > ey1) ? ey1 : y;
        if (x < ex0) x = ex0;
        else if (x > ex1) x = ex1;
        int64_t px = ((int64_t)(y - ey0) * L->sft_nx + (x - ex0)) * 3;
        add64(&T->sft[px], q0);
        add64(&T->sft[px + 1], q1);
        add64(&T->sft[px + 2], q2);
    }
    if (T->rtd && L->n_rtd) {
        int32_t ia = orc_cvt(((double)(r8[6] * 180.0f) / 3.14159265358979323846) / (double)P->det_grid[5]);
        int32_t ir = orc_cvt(r8[7] / P->det_grid[3]);
        if (ir > L->dimR - 1) ir = L->dimR - 1;
        else if (ir < 0) ir = 0;
        if (ia > L->dimAlpha - 1) ia = L->dimAlpha - 1;
        else if (ia < 0) ia = 0;
        int64_t px = ((int64_t)ia * L->dimR + ir) * 3;
        add64(&T->rtd[px], q0);
        add64(&T->rtd[px + 1], q1);
        add64(&T->rtd[px + 2], q2);
    }
    if (T->rad && L->n_rad) {
        int32_t ib = orc_cvt(r8[7] / P->rad_dr);
        if (ib >= 0 && ib < P->rad_n) {
            add64(&T->rad[3 * (int64_t)ib], q0);
            add64(&T->rad[3 * (int64_t)ib + 1], q1);
            add64(&T->rad[3 * (int64_t)ib + 2], q2);
        }
    }
    if (T->det) {
        for (int32_t k = 0; k < P->ndet; ++k) {
            if (r8[7] >= P->det_rlo[k] && r8[7] < P->det_rhi[k]) {
                add64(&T->det[2 * k], q0);
                add64(&T->det[2 * k + 1], q1);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Adjoint bookkeeping of one photon                                     */
/* ------------------------------------------------------------------ */
typedef struct {
    float *L;       /* path length per tissue, accumulated in step order */
    int32_t *n;     /* interactions per tissue */
    float *seg;     /* 8 floats per move: start (3), end (3) in the emitter frame, length,
                       and the tissue of the interaction that ends it (-1: a face) */
    int64_t nseg, cap;
} adj_track;

static void adj_move(adj_track *A, int32_t tissue, of3 c0, of3 c1, float s, int collision)
{
    A->L[tissue] += s;
    if (collision) A->n[tissue]++;
    if (A->nseg == A->cap) {
        A->cap = A->cap ? 2 * A->cap : 1024;
        A->seg = (float *)realloc(A->seg, (size_t)A->cap * 8 * sizeof(float));
    }
    float *q = A->seg + 8 * A->nseg++;
    q[0] = c0.x; q[1] = c0.y; q[2] = c0.z; q[3] = c1.x; q[4] = c1.y; q[5] = c1.z; q[6] = s;
    q[7] = collision ? (float)tissue : -1.0f;
}

/* Adds -(length in voxel) * W of the segment c0 -> c1 (length s) to each
 * voxel it crosses: Amanatides & Woo's traversal in grid units, clipped to
 * the grid; one axis advanced per crossing (x before y before z on ties). */
static void dda_add(const orc_adjoint_out *A, int64_t *jac_d, const float *q, float W)
{
    const int32_t *n = A->jac_n;
    float g0[3], d[3];
    for (int a = 0; a < 3; ++a) {
        float inv = 1.0f / A->jac_size[a];
        g0[a] = (q[a] - A->jac_origin[a]) * inv;
        d[a] = (q[3 + a] - A->jac_origin[a]) * inv - g0[a];
    }
    float t0 = 0.0f, t1 = 1.0f;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0f) {
            if (!(g0[a] >= 0.0f && g0[a] < (float)n[a])) return;
        } else {
            float ta = (0.0f - g0[a]) / d[a], tb = ((float)n[a] - g0[a]) / d[a];
            if (ta > tb) { float x = ta; ta = tb; tb = x; }
            if (ta > t0) t0 = ta;
            if (tb < t1) t1 = tb;
        }
    }
    if (!(t0 < t1)) return;
    int32_t i[3], st[3];
    float tmax[3], tdel[3];
    for (int a = 0; a < 3; ++a) {
        int32_t k = (int32_t)floorf(g0[a] + t0 * d[a]);
        if (k < 0) k = 0;
        if (k > n[a] - 1) k = n[a] - 1;
        i[a] = k;
        if (d[a] > 0.0f) { st[a] = 1; tmax[a] = ((float)(k + 1) - g0[a]) / d[a]; tdel[a] = 1.0f / d[a]; }
        else if (d[a] < 0.0f) { st[a] = -1; tmax[a] = ((float)k - g0[a]) / d[a]; tdel[a] = -1.0f / d[a]; }
        else { st[a] = 0; tmax[a] = INFINITY; tdel[a] = INFINITY; }
    }
    const float s = q[6];
    float t = t0;
    for (int32_t it = 0; it < n[0] + n[1] + n[2] + 3; ++it) {
        float tn = fminf(fminf(tmax[0], tmax[1]), fminf(tmax[2], t1));
        if (tn > t) {
            int64_t v = ((int64_t)i[2] * n[1] + i[1]) * n[0] + i[0];
            add64(&jac_d[v], fx32(-(((tn - t) * s) * W)));
            t = tn;
        }
        if (t >= t1) break;
        int a = (tmax[0] <= tmax[1] && tmax[0] <= tmax[2]) ? 0 : (tmax[1] <= tmax[2] ? 1 : 2);
        i[a] += st[a];
        if (i[a] < 0 || i[a] >= n[a]) break;
        tmax[a] += tdel[a];
    }
}

/* The interaction term of the per-voxel mus gradient: w = W / mus added to
 * the voxel holding the interaction point c (emitter frame), if any. */
static void coll_add(const orc_adjoint_out *A, int64_t *jac_d, const float *c, float w)
{
    const int32_t *n = A->jac_n;
    float g[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = (c[a] - A->jac_origin[a]) * (1.0f / A->jac_size[a]);
        if (!(g[a] >= 0.0f && g[a] < (float)n[a])) return;
    }
    add64(&jac_d[((int64_t)(int32_t)g[2] * n[1] + (int32_t)g[1]) * n[0] + (int32_t)g[0]], fx32(w));
}

/* TransmitOrReflect (.cl:921-1018) with GetCosAlphaCrit (:909-919),
 * GetFresnel (:857-907), Transmit (:506-529) and Reflect (:532-549): the
 * direction after a boundary event at unit normal nv between index n1 (the
 * photon's tissue) and n2, for the Fresnel draw x.  cos_in = nv . v is
 * signed as in the reference.  Returns ORC_EV_TRANSMIT or ORC_EV_REFLECT.
 * ORC_PLANT_* (test builds only) plant one mistake each, so the tests can
 * show that their checks catch it. */
static int transmit_or_reflect(of3 *vp, of3 nv, float n1, float n2, float x)
{
    const of3 v = *vp;
    const float cos_in = vdot(nv, v);
    const float abs_cos = fabsf(cos_in);
#ifdef ORC_PLANT_SWAP_N
    { float q = n1; n1 = n2; n2 = q; }
#endif
    float cos_crit = 0.0f;
    if (n1 > n2) { float q = n2 / n1; cos_crit = sqrtf(1.0f - q * q); }
    const float r = (abs_cos <= cos_crit) ? 1.0f : orc_fresnel(n1, n2, cos_in);
    const of3 ey = vnormalize(vcross(nv, vcross(nv, v)));
    if (x > r) { /* Transmit (.cl:506-529) */
        float sin_t = sqrtf(orc_clamp0(1 - cos_in * cos_in)) * n1 / n2;
#ifdef ORC_PLANT_FLIP_SIN_T
        sin_t = -sin_t;
#endif
        float cos_t = -sqrtf(orc_clamp0(1 - sin_t * sin_t));
        if (cos_in > 0.0f) cos_t = fabsf(cos_t);
        *vp = vnormalize(vsub(vmul(cos_t, nv), vmul(sin_t, ey)));
        return ORC_EV_TRANSMIT;
    }
    /* Reflect (.cl:532-549) */
    float sin_r = sqrtf(orc_clamp0(1 - cos_in * cos_in));
#ifdef ORC_PLANT_FLIP_EY
    sin_r = -sin_r;
#endif
    *vp = vneg(vnormalize(vadd(vmul(cos_in, nv), vmul(sin_r, ey))));
    return ORC_EV_REFLECT;
}

int orc_transmit_or_reflect(const float nv[3], float v[3], float n1, float n2, float x)
{
    of3 w = V(v[0], v[1], v[2]);
    int kind = transmit_or_reflect(&w, V(nv[0], nv[1], nv[2]), n1, n2, x);
    v[0] = w.x; v[1] = w.y; v[2] = w.z;
    return kind;
}

/* ------------------------------------------------------------------ */
/* One photon: kernel body .cl:1299-1598                                */
/* ------------------------------------------------------------------ */
static void run_photon(const orc_params *P, const orc_layout *L, int64_t id, orc_tallies *T,
                       int64_t *reg_loc, int64_t *misc_loc, float *exit8, float *trace,
                       int32_t trace_max, int32_t *trace_n, adj_track *A)
{
    const float *pb = P->poi_basis;
    const of3 POI = V(pb[0], pb[1], pb[2]);
    const of3 eX = V(pb[3], pb[4], pb[5]);
    const of3 eY = V(pb[6], pb[7], pb[8]);
    const of3 eZ = V(pb[9], pb[10], pb[11]);
    const float Rsp = L->rsp;
    const float distMin = L->dist_min;
    const float *NM = P->nmuasg;

    orc_rng rng;
    rng_init(&rng, P->rng_mode, P->seed, id);

    /* GetvStart (.cl:423-467): A == 1 takes vStart[0..2]; for A > 1 the
     * direction is chosen by the cumulative vStartOcc counts. */
    int32_t dsel = 0;
    if (P->nstart > 1) {
        int64_t sum = 0;
        for (int32_t i = 0; i < P->nstart; ++i) {
            sum += P->vocc[i];
            if (id < sum) { dsel = i; break; }
            dsel = i;
        }
    }
    /* PhotonClass::Init (.cl:344-389) */
    of3 v = vload3(P->vstart, dsel);
    of3 p = V(P->pstart[0], P->pstart[1], P->pstart[2]);
    float w = 1.0f - Rsp;
    of3 sft = V(0.0f, 0.0f, 0.0f);
    float stepsize = 0.0f, stb = 0.0f, unfin = 0.0f, absorbed = 0.0f;
    int fetal = 0;
    int32_t refTri = -1;
    int32_t curr = 0;
    of3 nv = V(0.0f, 0.0f, -1.0f);
    int roulette = 1;
    int64_t steps = 0;
    int64_t nbnd = 0; /* boundary events since the last interaction */
    int32_t ntr = 0;
    int64_t arz_out = 0, vox_out = 0;

    do {
        int kind;
        float pp[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z};
        int32_t tri = orc_find_triangle(P->tri, P->ntri, pp, vv, refTri, &stb);
        float c_n = NM[4 * curr], c_a = NM[4 * curr + 1], c_s = NM[4 * curr + 2], c_g = NM[4 * curr + 3];
        /* step length (.cl:1405-1442) */
        if (c_a == 0.0f && c_s == 0.0f) {
            stepsize = stb;
        } else if (unfin == 0.0f) {
            float x;
            do { x = rng_next(&rng); } while (x <= 0.0f);
            stepsize = -orc_logf(x) * P->factor / (c_a + c_s);
        } else {
            stepsize = unfin / (c_a + c_s);
            unfin = 0.0f;
        }
        if (stb <= stepsize) {
            /* boundary branch (.cl:1462-1511) */
            of3 p_old = p;
            p = vadd(p, vmul(stb, v)); /* MoveStep */
            if (A) adj_move(A, curr, calc_coords(p_old, POI, eX, eY, eZ), calc_coords(p, POI, eX, eY, eZ), stb, 0);
            refTri = -1;
            if (tri < 0) { /* no face ahead: the photon has left the mesh */
                kind = ORC_EV_ESCAPE;
                w = 0.0f;
                misc_loc[ORC_MISC_ESCAPED]++;
            } else {
                float x = rng_next(&rng);
                int32_t in_id = P->med[2 * tri], out_id = P->med[2 * tri + 1];
                nv = vnormalize(vload3(P->nrm, tri));
                int32_t nx_id = (curr == out_id) ? in_id : out_id;
                float n_n = NM[4 * nx_id], n_a = NM[4 * nx_id + 1], n_s = NM[4 * nx_id + 2];
                if (fabsf(sqrtf(p.x * p.x + p.y * p.y)) < 0.01f)
                    nv = vmul(orc_sign(vdot(nv, eZ)), vnormalize(eZ));
                /* TransmitOrReflect (.cl:921-1018) */
                const float cos_in = vdot(nv, v);
                int32_t new_id;
                float m_a, m_s;
                kind = transmit_or_reflect(&v, nv, c_n, n_n, x);
                if (kind == ORC_EV_TRANSMIT) {
                    new_id = (cos_in > 0) ? out_id : in_id;
                    m_a = n_a; m_s = n_s;
                } else {
                    new_id = curr;
                    m_a = c_a; m_s = c_s;
                }
                if (m_a == 0 || m_s == 0) { /* leaves the tissue: record and kill */
                    sft.y = 0;
                    float dist = fabsf(vdot(vnormalize(eZ), vsub(p, POI)));
                    if (dist > distMin) {
                        sft.z = w * (1 - Rsp);
                    } else {
                        if (fetal) sft.y = w * (1 - Rsp);
                        sft.x = w * (1 - Rsp);
                    }
                    w = 0.0f;
                }
                curr = new_id;
                refTri = tri;
                if (curr == P->fetal_id) fetal = 1;
                unfin = (stepsize - stb) * (c_a + c_s);
            }
            misc_loc[ORC_MISC_BOUNDARY]++;
        } else {
            /* interaction branch (.cl:1513-1537) */
            of3 p_old = p;
            p = vadd(p, vmul(stepsize, v));
            if (A) adj_move(A, curr, calc_coords(p_old, POI, eX, eY, eZ), calc_coords(p, POI, eX, eY, eZ), stepsize, 1);
            refTri = -1;
            if (c_a > 0.0f && c_s > 0.0f) { /* Absorb (.cl:841-847) */
                absorbed = w * (c_a / (c_a + c_s));
                w -= absorbed;
            }
            float xt = rng_next(&rng);
            float xp = rng_next(&rng);
            v = vscatter(v, c_g, xt, xp);
            of3 c = calc_coords(p, POI, eX, eY, eZ);
            int64_t loc[3] = {0, 0, 0};
            record_absorption(P, L, T, curr, absorbed, c_a, c, loc);
            reg_loc[curr] += loc[0];
            arz_out += loc[1];
            vox_out += loc[2];
            misc_loc[ORC_MISC_INTERACT]++;
            kind = ORC_EV_INTERACT;
        }
        /* Roulette (.cl:1020-1035, 1539-1545) */
        if (w < P->w_th && w > 0.0f) {
            float x = rng_next(&rng);
            if (x < P->chance) { w /= P->chance; roulette = 1; }
            else { w = 0; roulette = 0; }
        }
        steps++;
        if (trace && ntr < trace_max) {
            float *t = trace + 12 * (int64_t)ntr;
            t[0] = (float)kind; t[1] = (float)(kind == ORC_EV_INTERACT ? -1 : tri);
            t[2] = (float)curr; t[3] = w;
            t[4] = p.x; t[5] = p.y; t[6] = p.z; t[7] = v.x; t[8] = v.y; t[9] = v.z;
            t[10] = stepsize; t[11] = (float)rng.ndraw;
            ntr++;
        }
        /* Step cap, and the boundary-run cap: a photon caught on the edge of
         * two faces reflects from one to the other at t ~ 0 for ever (the
         * last-face exclusion cannot separate them; the reference loop never
         * ends on it).  No reference counterpart. */
        nbnd = kind == ORC_EV_INTERACT ? 0 : nbnd + 1;
        if (w > 0.0f && ((P->max_steps > 0 && steps >= P->max_steps) ||
                         (P->max_boundary_run > 0 && nbnd >= P->max_boundary_run))) {
            misc_loc[ORC_MISC_CAPPED]++;
            w = 0.0f;
        }
    } while (w > 0.0f && (w > P->w_th || roulette));
    if (p.x != p.x || p.y != p.y || p.z != p.z || v.x != v.x || v.y != v.y || v.z != v.z)
        misc_loc[ORC_MISC_NAN]++;
    misc_loc[ORC_MISC_STEPS] += steps;
    misc_loc[ORC_MISC_ARZ_OUT] += arz_out;
    misc_loc[ORC_MISC_VOX_OUT] += vox_out;

    /* epilogue (.cl:1586-1594) */
    float r8[8];
    of3 c = calc_coords(p, POI, eX, eY, eZ);
    of3 dn = vsub(POI, p);
    r8[0] = c.x; r8[1] = c.y; r8[2] = c.z;
    r8[3] = sft.x; r8[4] = sft.y; r8[5] = sft.z;
    r8[6] = fabsf(orc_acosf(vdot(v, nv)));
    r8[7] = sqrtf((dn.x * dn.x + dn.y * dn.y) + dn.z * dn.z);
    if (exit8) memcpy(exit8, r8, sizeof r8);
    if (trace_n) *trace_n = ntr;
    record_exit(P, L, T, r8);
    if (A) { /* R weight and exit radius for the detector test */
        A->L[P->ntissue] = r8[3];
        A->L[P->ntissue + 1] = r8[7];
    }
}

int orc_run(const orc_params *P, int64_t first, int64_t count, orc_tallies *T, float *exit_rec,
            float *trace, int32_t trace_max, int32_t *trace_n, int nthreads)
{
    if (!P || !T || count < 0 || P->ntissue < 2 || P->nstart < 1) return 1;
    orc_layout L;
    orc_layout_compute(P, &L);
    int64_t misc[ORC_MISC_N] = {0};
    int64_t *reg = (int64_t *)calloc((size_t)P->ntissue, sizeof(int64_t));
    if (!reg) return 2;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int64_t misc_l[ORC_MISC_N] = {0};
        int64_t *reg_l = (int64_t *)calloc((size_t)P->ntissue, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < count; ++i) {
            int64_t id = first + i;
            run_photon(P, &L, id, T, reg_l, misc_l, exit_rec ? exit_rec + 8 * i : NULL,
                       trace ? trace + 12 * (int64_t)trace_max * i : NULL, trace_max,
                       trace_n ? trace_n + i : NULL, NULL);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int k = 0; k < ORC_MISC_N; ++k) misc[k] += misc_l[k];
            for (int32_t k = 0; k < P->ntissue; ++k) reg[k] += reg_l[k];
        }
        free(reg_l);
    }
    misc[ORC_MISC_PHOTONS] += count;
    for (int k = 0; k < ORC_MISC_N; ++k) T->misc[k] += misc[k];
    if (T->region)
        for (int32_t k = 0; k < P->ntissue; ++k) T->region[k] += reg[k];
    free(reg);
    return 0;
}

int orc_adjoint(const orc_params *P, int64_t first, int64_t count, orc_tallies *T, orc_adjoint_out *out,
                int nthreads)
{
    if (!P || !out || count < 0 || P->ntissue < 2 || P->nstart < 1) return 1;
    /* the path-length estimator assumes step lengths -ln(xi) / mu_t */
    if (P->factor != 1.0f) return 3;
    orc_layout L;
    orc_layout_compute(P, &L);
    const int32_t nt = P->ntissue;
    const int64_t nvox = (int64_t)out->jac_n[0] * out->jac_n[1] * out->jac_n[2];
    int64_t misc[ORC_MISC_N] = {0}, pairs = 0;
    int64_t *reg = (int64_t *)calloc((size_t)nt, sizeof(int64_t));
    if (!reg) return 2;
    orc_tallies none;
    memset(&none, 0, sizeof none);
    orc_tallies *TT = T ? T : &none;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : pairs)
#endif
    {
        int64_t misc_l[ORC_MISC_N] = {0};
        int64_t *reg_l = (int64_t *)calloc((size_t)nt, sizeof(int64_t));
        adj_track A;
        A.L = (float *)calloc((size_t)nt + 2, sizeof(float)); /* + R weight, exit radius */
        A.n = (int32_t *)calloc((size_t)nt, sizeof(int32_t));
        A.seg = NULL;
        A.cap = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < count; ++i) {
            memset(A.L, 0, ((size_t)nt + 2) * sizeof(float));
            memset(A.n, 0, (size_t)nt * sizeof(int32_t));
            A.nseg = 0;
            run_photon(P, &L, first + i, TT, reg_l, misc_l, NULL, NULL, 0, NULL, &A);
            const float W = A.L[nt], radius = A.L[nt + 1];
            if (out->path_len) {
                memcpy(out->path_len + i * nt, A.L, (size_t)nt * sizeof(float));
                memcpy(out->path_int + i * nt, A.n, (size_t)nt * sizeof(int32_t));
                out->path_w[i] = W;
                out->path_r[i] = radius;
            }
            if (W == 0.0f) continue;
            for (int32_t d = 0; d < P->ndet; ++d) {
                if (!(radius >= P->det_rlo[d] && radius < P->det_rhi[d])) continue;
                pairs++;
                if (out->signals) add64(&out->signals[d], fx32(W));
                if (out->grad)
                    for (int32_t t = 0; t < nt; ++t) {
                        if (A.L[t] == 0.0f && A.n[t] == 0) continue;
                        float mus = P->nmuasg[4 * t + 2];
                        float gm = -(A.L[t] * W);
                        float gs = mus != 0.0f ? (((float)A.n[t] / mus) - A.L[t]) * W : gm;
                        add64(&out->grad[((int64_t)d * nt + t) * 2], fx32(gm));
                        add64(&out->grad[((int64_t)d * nt + t) * 2 + 1], fx32(gs));
                    }
                if (out->jac && nvox > 0)
                    for (int64_t k = 0; k < A.nseg; ++k) dda_add(out, out->jac + d * nvox, A.seg + 8 * k, W);
                if (out->jac_mus && nvox > 0)
                    for (int64_t k = 0; k < A.nseg; ++k) {
                        const float *q = A.seg + 8 * k;
                        dda_add(out, out->jac_mus + d * nvox, q, W);
                        const float mus = q[7] >= 0.0f ? P->nmuasg[4 * (int32_t)q[7] + 2] : 0.0f;
                        if (mus != 0.0f) coll_add(out, out->jac_mus + d * nvox, q + 3, W / mus);
                    }
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int k = 0; k < ORC_MISC_N; ++k) misc[k] += misc_l[k];
            for (int32_t k = 0; k < nt; ++k) reg[k] += reg_l[k];
        }
        free(reg_l);
        free(A.L);
        free(A.n);
        free(A.seg);
    }
    out->ndetected += pairs;
    if (T) {
        misc[ORC_MISC_PHOTONS] += count;
        for (int k = 0; k < ORC_MISC_N; ++k) T->misc[k] += misc[k];
        if (T->region)
            for (int32_t k = 0; k < nt; ++k) T->region[k] += reg[k];
    }
    free(reg);
    return 0;
}

void orc_dda(const orc_adjoint_out *out, const float seg[7], float w, int64_t *jac) { dda_add(out, jac, seg, w); }
