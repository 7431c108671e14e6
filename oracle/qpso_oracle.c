/*
 * qpso_oracle.c -- CPU restatement of the Q-PSO hot path (see qpso_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fwrapv [-fopenmp].
 * Every citation is file:line in the reference repository.
 */
#include "qpso_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* wrapping 32-bit integer helpers (the reference's int math wraps on HW)   */
/* ------------------------------------------------------------------------ */
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static inline int32_t wshl(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* __float2int_rn (main.h:1331-1338, 1447): round half to even, saturating, NaN -> 0. */
int32_t or_rint_sat(float x)
{
    if (isnan(x)) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)rintf(x);
}

/* __frcp_rn: correctly rounded reciprocal. */
float or_rcp(float x) { return 1.0f / x; }

/* ------------------------------------------------------------------------ */
/* sin / cos / acos: own implementations, evaluated in binary64 with only    */
/* IEEE +,-,*,fma,sqrt and rounded once to binary32.                        */
/*  sin/cos: Cody-Waite reduction by pi/2 (two-part constant) and Taylor     */
/*  series of sin to r^15 / cos to r^16 on |r| <= pi/4.                      */
/*  acos: acos(x) = pi/2 - asin(x) for |x| <= 1/2, and                       */
/*  acos(|x|) = 2 asin(sqrt((1-|x|)/2)) above; asin(s) = s + s z P(z),       */
/*  z = s^2, P the Taylor series of (asin(s)-s)/s^3 to z^19.                 */
/* ------------------------------------------------------------------------ */
static const double OR_PIO2_HI = 0x1.921fb54442d18p+0;
static const double OR_PIO2_LO = 0x1.1a62633145c07p-54;
static const double OR_PI = 0x1.921fb54442d18p+1;
static const double OR_2_OVER_PI = 0x1.45f306dc9c883p-1;

static const double OR_SIN_C[7] = {  /* (-1)^n / (2n+1)!, n = 1..7 */
    -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
    0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
    -0x1.ae7f3e733b81fp-41};
static const double OR_COS_C[8] = {  /* (-1)^n / (2n)!, n = 1..8 */
    -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
    0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
    -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45};
static const double OR_ASIN_C[20] = { /* (2n)! / (4^n (n!)^2 (2n+1)), n = 1..20 */
    0x1.5555555555555p-3, 0x1.3333333333333p-4, 0x1.6db6db6db6db7p-5,
    0x1.f1c71c71c71c7p-6, 0x1.6e8ba2e8ba2e9p-6, 0x1.1c4ec4ec4ec4fp-6,
    0x1.c99999999999ap-7, 0x1.7a87878787878p-7, 0x1.3fde50d79435ep-7,
    0x1.12ef3cf3cf3cfp-7, 0x1.df3bd37a6f4dfp-8, 0x1.a6863d70a3d71p-8,
    0x1.782dda12f684cp-8, 0x1.51ba308d3dcb1p-8, 0x1.31683bdef7bdfp-8,
    0x1.15ee9d45d1746p-8, 0x1.fcaf8fb6db6dbp-9, 0x1.d3d2a8e0dd67dp-9,
    0x1.b026f57b13b14p-9, 0x1.90cb77f60c7cep-9};

static double or_sin_poly(double r)
{
    double z = r * r;
    double p = OR_SIN_C[6];
    for (int i = 5; i >= 0; --i) p = fma(p, z, OR_SIN_C[i]);
    return fma(r * z, p, r);
}

static double or_cos_poly(double r)
{
    double z = r * r;
    double p = OR_COS_C[7];
    for (int i = 6; i >= 0; --i) p = fma(p, z, OR_COS_C[i]);
    return fma(z, p, 1.0);
}

/* Reduce x to r in [-pi/4, pi/4] (approximately) and quadrant q = k mod 4. */
static double or_reduce(double x, int* q)
{
    double k = rint(x * OR_2_OVER_PI);
    double r = fma(-k, OR_PIO2_HI, x);
    r = fma(-k, OR_PIO2_LO, r);
    double m = k - 4.0 * floor(k * 0.25); /* exact for |k| < 2^53 */
    *q = (int)m;
    return r;
}

float or_sinf(float xf)
{
    double x = (double)xf;
    if (!isfinite(x)) return (float)(x - x);
    int q;
    double r = or_reduce(x, &q);
    double s = or_sin_poly(r), c = or_cos_poly(r);
    double v = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
    return (float)v;
}

float or_cosf(float xf)
{
    double x = (double)xf;
    if (!isfinite(x)) return (float)(x - x);
    int q;
    double r = or_reduce(x, &q);
    double s = or_sin_poly(r), c = or_cos_poly(r);
    double v = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
    return (float)v;
}

static double or_asin_core(double s, double z)
{
    double p = OR_ASIN_C[19];
    for (int i = 18; i >= 0; --i) p = fma(p, z, OR_ASIN_C[i]);
    return fma(s * z, p, s);
}

float or_acosf(float xf)
{
    double x = (double)xf;
    if (isnan(x)) return xf;
    double ax = fabs(x);
    if (ax > 1.0) return (float)((x - x) / (x - x)); /* NaN, as IEEE acos */
    if (ax <= 0.5) {
        double a = or_asin_core(x, x * x);
        return (float)(OR_PIO2_HI - a);
    }
    double z = (1.0 - ax) * 0.5;
    double s = sqrt(z);
    double a2 = 2.0 * or_asin_core(s, z);
    return (float)(x > 0.0 ? a2 : OR_PI - a2);
}

/* ------------------------------------------------------------------------ */
/* random table (main.cu:1337-1352)                                         */
/* GSL's default generator is mt19937 with seed 0, and gsl_rng_uniform      */
/* returns (double)u32 / 2^32; GSL replaces seed 0 with 4357 and seeds the   */
/* state with the 2002 Matsumoto-Nishimura init_genrand recurrence.          */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t mt[624]; int mti; } or_mt;

static void or_mt_seed(or_mt* m, uint32_t s)
{
    if (s == 0) s = 4357u;
    m->mt[0] = s;
    for (int i = 1; i < 624; ++i)
        m->mt[i] = 1812433253u * (m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) + (uint32_t)i;
    m->mti = 624;
}

static uint32_t or_mt_next(or_mt* m)
{
    if (m->mti >= 624) {
        for (int kk = 0; kk < 624; ++kk) {
            uint32_t y = (m->mt[kk] & 0x80000000u) | (m->mt[(kk + 1) % 624] & 0x7fffffffu);
            m->mt[kk] = m->mt[(kk + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        m->mti = 0;
    }
    uint32_t y = m->mt[m->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void or_mt19937_raw(uint32_t seed, uint32_t* out, size_t n)
{
    or_mt m;
    or_mt_seed(&m, seed);
    for (size_t i = 0; i < n; ++i) out[i] = or_mt_next(&m);
}

void or_rand_table(uint32_t seed, float* out, size_t n)
{
    or_mt m;
    or_mt_seed(&m, seed);
    for (size_t i = 0; i < n; ++i) out[i] = (float)((double)or_mt_next(&m) / 4294967296.0);
}

size_t or_rand_stride(int n_iters)
{
    size_t need = (size_t)12 * (size_t)(n_iters + 1);
    return need > 512 ? need : 512;
}

/* ------------------------------------------------------------------------ */
/* quaternion helpers, state vector [tx,ty,tz,q0,q1,q2,q3] (main.h:86-94)    */
/* ------------------------------------------------------------------------ */
/* QuatProd_pqPSOstateVect (main.h:249-255): p (x) q, Hamilton, scalar first. */
static void qprod(const float p[4], const float q[4], float o[4])
{
    float a;
    a = fmaf(p[0], q[0], -(p[1] * q[1])); a = fmaf(-p[2], q[2], a); o[0] = fmaf(-p[3], q[3], a);
    a = fmaf(p[1], q[0], p[0] * q[1]);    a = fmaf(p[2], q[3], a);  o[1] = fmaf(-p[3], q[2], a);
    a = fmaf(p[2], q[0], p[0] * q[2]);    a = fmaf(p[3], q[1], a);  o[2] = fmaf(-p[1], q[3], a);
    a = fmaf(p[3], q[0], p[0] * q[3]);    a = fmaf(p[1], q[2], a);  o[3] = fmaf(-p[2], q[1], a);
}

/* QuatProd_pq_1PSOstateVect (main.h:257-264): p (x) conj(q). */
static void qprod_conj(const float p[4], const float q[4], float o[4])
{
    float a;
    a = fmaf(p[0], q[0], p[1] * q[1]);    a = fmaf(p[2], q[2], a);  o[0] = fmaf(p[3], q[3], a);
    a = fmaf(p[1], q[0], -(p[0] * q[1])); a = fmaf(-p[2], q[3], a); o[1] = fmaf(p[3], q[2], a);
    a = fmaf(p[2], q[0], -(p[0] * q[2])); a = fmaf(-p[3], q[1], a); o[2] = fmaf(p[1], q[3], a);
    a = fmaf(p[3], q[0], -(p[0] * q[3])); a = fmaf(-p[1], q[2], a); o[3] = fmaf(p[2], q[1], a);
}

/* Cast2TopHalfHypersphere (main.h:224-246). */
static void top_half(float q[4])
{
    if (q[0] < 0.0f) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; q[3] = -q[3]; }
}

/* NormalizeQuaternionPSOstateVect (main.h:210-221). */
static void normalize(float q[4])
{
    float s = fmaf(q[0], q[0], q[1] * q[1]);
    s = fmaf(q[2], q[2], s);
    s = fmaf(q[3], q[3], s);
    float d = sqrtf(s);
    if (d > 1.0e-7f) {
        float inv = or_rcp(d);
        q[0] *= inv; q[1] *= inv; q[2] *= inv; q[3] *= inv;
    }
}

/* QuatKinematicsPSOStateVect / ...WithInit (main.h:267-307, 374-413):
 * q <- normalize(top_half(cos(d) q + 1/2 sinc(d) q (x) (0, w/2))), d = |w/2|/2. */
static void kinematics_from(const float src[4], float dst[4], const float w[3])
{
    float wt[4] = {0.0f, w[0] * 0.5f, w[1] * 0.5f, w[2] * 0.5f}; /* * mTs_ (0.5) */
    float n2 = fmaf(wt[1], wt[1], wt[2] * wt[2]);
    n2 = fmaf(wt[3], wt[3], n2);
    float d = sqrtf(n2) * 0.5f;
    float cw = or_cosf(d);
    float sw = (d < 1.0e-7f) ? 1.0f : or_sinf(d) * or_rcp(d);
    float qxw[4];
    qprod(src, wt, qxw);
    float hs = 0.5f * sw;
    for (int i = 0; i < 4; ++i) dst[i] = fmaf(cw, src[i], hs * qxw[i]);
    normalize(dst);
    top_half(dst);
}

void or_kinematics(float q[4], const float w[3])
{
    float src[4] = {q[0], q[1], q[2], q[3]};
    kinematics_from(src, q, w);
}

/* ------------------------------------------------------------------------ */
/* projection (main.h:514-545, 309-369)                                     */
/* ------------------------------------------------------------------------ */
void or_project(const float pose[7], const float v[3], const or_cam* cam, float uvz[3])
{
    const float q0 = pose[3], q1 = pose[4], q2 = pose[5], q3 = pose[6];
    /* t = 2 * cross(q.xyz, v) */
    float t0 = 2.0f * fmaf(q2, v[2], -(q3 * v[1]));
    float t1 = 2.0f * -fmaf(q1, v[2], -(q3 * v[0]));
    float t2 = 2.0f * fmaf(q1, v[1], -(q2 * v[0]));
    /* k = cross(q.xyz, t) */
    float k0 = fmaf(q2, t2, -(q3 * t1));
    float k1 = -fmaf(q1, t2, -(q3 * t0));
    float k2 = fmaf(q1, t1, -(q2 * t0));
    /* v' = v + q0 t + k, then + translation */
    float px = fmaf(q0, t0, v[0]) + k0;
    float py = fmaf(q0, t1, v[1]) + k1;
    float pz = fmaf(q0, t2, v[2]) + k2;
    px += pose[0];
    py += pose[1];
    pz += pose[2];
    float r = or_rcp(pz);
    uvz[0] = (float)fma(cam->fx, (double)(px * r), cam->cx);
    uvz[1] = (float)fma(cam->fy, (double)(py * r), cam->cy);
    uvz[2] = 1000.0f * pz; /* tomm_ */
}

/* ------------------------------------------------------------------------ */
/* render (main.cu:305-419, main.h:1306-1491, AABB main.h:1190-1304 +        */
/* main.cu:423-546)                                                         */
/* ------------------------------------------------------------------------ */
#define OR_FXP_LIMIT (1 << 30)
#define OR_AREA_LIMIT ((int64_t)1 << 23)

typedef struct {
    int32_t X[3], Y[3];
    float Z[3];
    int32_t minx, maxx, miny, maxy;
} or_face;

static void face_project(const float* mesh, int F, int f, const float pose[7],
                         const or_cam* cam, or_face* fc, int* degen)
{
    for (int j = 0; j < 3; ++j) {
        /* loadOBJ3 SoA layout (main.h:620-679): x = A[f + j F], y = A[3F + f + jF], z = A[6F + ...] */
        float v[3] = {mesh[f + j * F], mesh[3 * F + f + j * F], mesh[6 * F + f + j * F]};
        float uvz[3];
        or_project(pose, v, cam, uvz);
        fc->X[j] = or_rint_sat(16.0f * uvz[0]);
        fc->Y[j] = or_rint_sat(16.0f * uvz[1]);
        fc->Z[j] = uvz[2];
        if (fc->X[j] > OR_FXP_LIMIT || fc->X[j] < -OR_FXP_LIMIT ||
            fc->Y[j] > OR_FXP_LIMIT || fc->Y[j] < -OR_FXP_LIMIT)
            *degen = 1;
    }
    /* (min + 0xF) >> 4; no wrap once |X| <= 2^30 */
    fc->minx = (imin(imin(fc->X[0], fc->X[1]), fc->X[2]) + 15) >> 4;
    fc->maxx = (imax(imax(fc->X[0], fc->X[1]), fc->X[2]) + 15) >> 4;
    fc->miny = (imin(imin(fc->Y[0], fc->Y[1]), fc->Y[2]) + 15) >> 4;
    fc->maxy = (imax(imax(fc->Y[0], fc->Y[1]), fc->Y[2]) + 15) >> 4;
}

/* Rasterize one face: "bbox fill".  The reference's inside test
 * (main.h:1440, 1473) is always true (mask is 0/1, never has bit 31), so every
 * pixel of the face's integer bbox [minx,maxx)x[miny,maxy) receives
 * atomicMin(rint(z1 + cx2*z2 + cx3*z3)).  Edge values are evaluated directly
 * at the first in-frame pixel; mod 2^32 this equals the reference's
 * incremental CX -= FDY / CY += FDX walk. */
static uint64_t face_raster(const or_face* fc, int32_t* frame)
{
    const int32_t X1 = fc->X[0], X2 = fc->X[1], X3 = fc->X[2];
    const int32_t Y1 = fc->Y[0], Y2 = fc->Y[1], Y3 = fc->Y[2];
    const int32_t DX12 = wsub(X1, X2), DX23 = wsub(X2, X3), DX31 = wsub(X3, X1);
    const int32_t DY12 = wsub(Y1, Y2), DY23 = wsub(Y2, Y3), DY31 = wsub(Y3, Y1);
    const int32_t tri = wsub(wmul(DX12, DY31), wmul(DX31, DY12));
    if (tri <= 0) return 0;
    if (fc->maxx - fc->minx <= 0 || fc->maxy - fc->miny <= 0) return 0;
    const float inv = 1.0f * or_rcp((float)tri); /* overA_ == 1.0f */
    const int32_t FDX23 = wshl(DX23, 4), FDX31 = wshl(DX31, 4);
    const int32_t FDY23 = wshl(DY23, 4), FDY31 = wshl(DY31, 4);
    int32_t C2 = wsub(wmul(DY23, X2), wmul(DX23, Y2));
    int32_t C3 = wsub(wmul(DY31, X3), wmul(DX31, Y3));
    C2 = wadd(C2, (DY23 < 0 || (DY23 == 0 && DX23 > 0)) ? 1 : 0);
    C3 = wadd(C3, (DY31 < 0 || (DY31 == 0 && DX31 > 0)) ? 1 : 0);
    const int32_t MINY = wshl(fc->miny, 4), MINX = wshl(fc->minx, 4);
    const int32_t CY2 = wsub(wadd(C2, wmul(DX23, MINY)), wmul(DY23, MINX));
    const int32_t CY3 = wsub(wadd(C3, wmul(DX31, MINY)), wmul(DY31, MINX));
    const float z1 = fc->Z[0];
    const float z2 = (fc->Z[1] - fc->Z[0]) * inv;
    const float z3 = (fc->Z[2] - fc->Z[0]) * inv;

    const int32_t x0 = imax(fc->minx, 0), x1 = imin(fc->maxx, OR_COLS);
    const int32_t y0 = imax(fc->miny, 0), y1 = imin(fc->maxy, OR_ROWS);
    if (x0 >= x1 || y0 >= y1) return 0; /* bbox entirely outside the frame */
    uint64_t n = 0;
    for (int32_t y = y0; y < y1; ++y) {
        int32_t cy2 = wadd(CY2, wmul(y - fc->miny, FDX23));
        int32_t cy3 = wadd(CY3, wmul(y - fc->miny, FDX31));
        int32_t CX2 = wsub(cy2, wmul(x0 - fc->minx, FDY23));
        int32_t CX3 = wsub(cy3, wmul(x0 - fc->minx, FDY31));
        int32_t* row = frame + (size_t)y * OR_COLS;
        for (int32_t x = x0; x < x1; ++x) {
            const int32_t cx2 = wadd(CX2, 15) >> 8;
            const int32_t cx3 = wadd(CX3, 15) >> 8;
            float d = fmaf((float)cx2, z2, z1);
            d = fmaf((float)cx3, z3, d);
            const int32_t di = or_rint_sat(d);
            if (di < row[x]) row[x] = di;
            CX2 = wsub(CX2, FDY23);
            CX3 = wsub(CX3, FDY31);
        }
        n += (uint64_t)(x1 - x0);
    }
    return n;
}

/* The same in the reference-faithful off-frame mode: the whole bbox is
 * walked as main.h:1424-1488 does and each pixel goes to the linear index
 * idx = y*cols_ + x (main.h:1442), in wrapping 32-bit arithmetic; an index in
 * [0, 640*480) is the particle's own buffer, any other one belongs to another
 * particle or lies outside the allocation and is dropped. */
static uint64_t face_raster_wrap(const or_face* fc, int32_t* frame)
{
    const int32_t X1 = fc->X[0], X2 = fc->X[1], X3 = fc->X[2];
    const int32_t Y1 = fc->Y[0], Y2 = fc->Y[1], Y3 = fc->Y[2];
    const int32_t DX12 = wsub(X1, X2), DX23 = wsub(X2, X3), DX31 = wsub(X3, X1);
    const int32_t DY12 = wsub(Y1, Y2), DY23 = wsub(Y2, Y3), DY31 = wsub(Y3, Y1);
    const int32_t tri = wsub(wmul(DX12, DY31), wmul(DX31, DY12));
    if (tri <= 0) return 0;
    if (fc->maxx - fc->minx <= 0 || fc->maxy - fc->miny <= 0) return 0;
    const float inv = 1.0f * or_rcp((float)tri);
    const int32_t FDX23 = wshl(DX23, 4), FDX31 = wshl(DX31, 4);
    const int32_t FDY23 = wshl(DY23, 4), FDY31 = wshl(DY31, 4);
    int32_t C2 = wsub(wmul(DY23, X2), wmul(DX23, Y2));
    int32_t C3 = wsub(wmul(DY31, X3), wmul(DX31, Y3));
    C2 = wadd(C2, (DY23 < 0 || (DY23 == 0 && DX23 > 0)) ? 1 : 0);
    C3 = wadd(C3, (DY31 < 0 || (DY31 == 0 && DX31 > 0)) ? 1 : 0);
    const int32_t MINY = wshl(fc->miny, 4), MINX = wshl(fc->minx, 4);
    int32_t CY2 = wsub(wadd(C2, wmul(DX23, MINY)), wmul(DY23, MINX));
    int32_t CY3 = wsub(wadd(C3, wmul(DX31, MINY)), wmul(DY31, MINX));
    const float z1 = fc->Z[0];
    const float z2 = (fc->Z[1] - fc->Z[0]) * inv;
    const float z3 = (fc->Z[2] - fc->Z[0]) * inv;
    uint64_t n = 0;
    for (int32_t y = fc->miny; y < fc->maxy; ++y) {
        int32_t CX2 = CY2, CX3 = CY3;
        for (int32_t x = fc->minx; x < fc->maxx; ++x) {
            const int32_t idx = wadd(wmul(y, OR_COLS), x);
            if (idx >= 0 && idx < OR_FRAME) {
                const int32_t cx2 = wadd(CX2, 15) >> 8;
                const int32_t cx3 = wadd(CX3, 15) >> 8;
                float d = fmaf((float)cx2, z2, z1);
                d = fmaf((float)cx3, z3, d);
                const int32_t di = or_rint_sat(d);
                if (di < frame[idx]) frame[idx] = di;
                ++n;
            }
            CX2 = wsub(CX2, FDY23);
            CX3 = wsub(CX3, FDY31);
        }
        CY2 = wadd(CY2, FDX23);
        CY3 = wadd(CY3, FDX31);
    }
    return n;
}

static int union_degenerate(const int32_t aabb[4], const int rect[4])
{
    const int32_t minX = imin(aabb[0], rect[0]), maxX = imax(aabb[1], rect[2]);
    const int32_t minY = imin(aabb[2], rect[1]), maxY = imax(aabb[3], rect[3]);
    const int64_t W = (int64_t)maxX - minX, H = (int64_t)maxY - minY;
    if (W <= 0 || H <= 0) return 1;
    return W * H > OR_AREA_LIMIT;
}

void or_render(const float* mesh, int F, int face_cap, const float pose[7],
               const or_cam* cam, const int rect[4], int32_t* frame, or_render_info* info)
{
    or_render_ex(mesh, F, face_cap, pose, cam, rect, 0, frame, info);
}

void or_render_ex(const float* mesh, int F, int face_cap, const float pose[7], const or_cam* cam,
                  const int rect[4], int wrap, int32_t* frame, or_render_info* info)
{
    const int Fe = (face_cap > 0 && face_cap < F) ? face_cap : F;
    /* sentinels of the AABB partial buffers (main.cu:1517-1535) take part in the
     * min/max (main.h:1279-1282, main.cu:496-531) */
    int32_t a0 = OR_AABB_MIN_SENT, a1 = OR_AABB_MAX_SENT;
    int32_t a2 = OR_AABB_MIN_SENT, a3 = OR_AABB_MAX_SENT;
    int degen = 0;
    or_face* faces = (or_face*)malloc(sizeof(or_face) * (size_t)(Fe > 0 ? Fe : 1));
    for (int f = 0; f < Fe; ++f) {
        face_project(mesh, F, f, pose, cam, &faces[f], &degen);
        a0 = imin(a0, faces[f].minx);
        a1 = imax(a1, faces[f].maxx);
        a2 = imin(a2, faces[f].miny);
        a3 = imax(a3, faces[f].maxy);
    }
    info->aabb[0] = a0; info->aabb[1] = a1; info->aabb[2] = a2; info->aabb[3] = a3;
    if (!degen) degen = union_degenerate(info->aabb, rect);
    info->degenerate = degen;
    info->raster_px = 0;
    if (!degen)
        for (int f = 0; f < Fe; ++f)
            info->raster_px += wrap ? face_raster_wrap(&faces[f], frame) : face_raster(&faces[f], frame);
    free(faces);
}

/* ------------------------------------------------------------------------ */
/* fitness (main.cu:550-764, main.h:1522-1573)                               */
/* ------------------------------------------------------------------------ */
float or_fitness(const or_render_info* rinfo, const int rect[4], uint32_t num_points,
                 int32_t* frame, const uint16_t* obs, or_fitness_info* finfo)
{
    return or_fitness_ex(rinfo, rect, num_points, 0, frame, obs, finfo);
}

float or_fitness_ex(const or_render_info* rinfo, const int rect[4], uint32_t num_points, int wrap,
                    int32_t* frame, const uint16_t* obs, or_fitness_info* finfo)
{
    memset(finfo, 0, sizeof(*finfo));
    if (rinfo->degenerate) {
        for (size_t i = 0; i < OR_FRAME; ++i) frame[i] = OR_EMPTY;
        finfo->err = NAN;
        return NAN;
    }
    const int32_t minX = imin(rinfo->aabb[0], rect[0]), maxX = imax(rinfo->aabb[1], rect[2]);
    const int32_t minY = imin(rinfo->aabb[2], rect[1]), maxY = imax(rinfo->aabb[3], rect[3]);
    const int32_t W = maxX - minX, H = maxY - minY, A = W * H;
    /* NumIters = __float2int_ru(Area * rcp(1024)) (main.cu:621-623); exact here */
    const int32_t nIt = (int32_t)ceilf((float)A * or_rcp(1024.0f));
    const int32_t T = nIt * 1024;
    const float rW = or_rcp((float)W);

    static _Thread_local float err[2048];
    static _Thread_local uint32_t seg[2048], tr[2048];
    static _Thread_local int32_t* visited = NULL;
    static _Thread_local size_t visited_cap = 0;
    memset(err, 0, sizeof(err));
    memset(seg, 0, sizeof(seg));
    memset(tr, 0, sizeof(tr));
    uint32_t nR = 0;
    size_t nv = 0;

    /* Thread tIdx visits t = tIdx + 1024*xx, xx = 0..nIt-1 (main.cu:632-733); each
     * slot t & 2047 therefore accumulates its pixels in increasing t.  A single
     * ascending scan over t reproduces every slot's order. */
    for (int32_t t = 0; t < T; ++t) {
        const int32_t xb = t % W;
        const int32_t yb = (int32_t)truncf((float)t * rW); /* __float2int_rz(t*rcp(W)) */
        const int32_t px = minX + xb, py = minY + yb;
        if (py > maxY) break; /* yb is non-decreasing in t */
        size_t idx;
        if (wrap) {
            /* pxIdxRendered = pxY*cols_ + pxX (+ the particle's offset) and
             * pxIdxKinect = pxY*cols_ + pxX (main.cu:674, 688), 32-bit */
            const int32_t L = wadd(wmul(py, OR_COLS), px);
            if (L < 0 || L >= OR_FRAME) continue; /* another buffer: read as empty */
            idx = (size_t)L;
        } else {
            if (px < 0 || px >= OR_COLS || py < 0 || py >= OR_ROWS) continue; /* Q4: empty */
            idx = (size_t)py * OR_COLS + (size_t)px;
        }
        const int32_t zr = frame[idx];
        const int32_t zk = (int32_t)obs[idx];
        const int s = t & 2047;
        if (nv == visited_cap) {
            visited_cap = visited_cap ? visited_cap * 2 : 65536;
            visited = (int32_t*)realloc(visited, visited_cap * sizeof(int32_t));
        }
        visited[nv++] = (int32_t)idx;
        if (zk != 0 && zr == OR_EMPTY) seg[s] += 1;
        if (zr != OR_EMPTY) {
            nR += 1;
            if (zk == 0) {
                tr[s] += 1;
            } else {
                const float e = (float)wsub(zr, zk) * 0.001f; /* tom_ */
                err[s] = fmaf(e, e, err[s]);
            }
        }
    }
    /* Q3: a pixel visited twice reads the rendered value both times; clear after. */
    for (size_t i = 0; i < nv; ++i) frame[visited[i]] = OR_EMPTY;

    /* FitnessReductionBallot halving tree (main.h:1530-1567) */
    for (int stride = 1024; stride >= 1; stride >>= 1)
        for (int i = 0; i < stride; ++i) err[i] = err[i] + err[i + stride];
    uint32_t sg = 0, st = 0;
    for (int i = 0; i < 2048; ++i) { sg += seg[i]; st += tr[i]; }

    const float inv_r = or_rcp((float)nR);
    const float ratio = (float)sg * or_rcp((float)num_points);
    const float ratioT = (float)st * inv_r;
    const float zerr = err[0] * inv_r;
    finfo->seg = sg; finfo->tr = st; finfo->n_rendered = nR; finfo->err = err[0];
    finfo->visit_px = nv;
    return (float)fma((double)(ratio + ratioT), 0.05, (double)zerr); /* main.cu:760 */
}

/* ------------------------------------------------------------------------ */
/* argmin (main.cu:1691-1698), Q8: NaN ranks above everything, first index   */
/* ------------------------------------------------------------------------ */
int or_argmin(const float* fit, int n)
{
    int best = 0;
    float bv = fit[0];
    for (int i = 1; i < n; ++i) {
        const float v = fit[i];
        if (isnan(v)) continue;
        if (isnan(bv) || v < bv) { bv = v; best = i; }
    }
    return best;
}

/* ------------------------------------------------------------------------ */
/* init (main.cu:72-198)                                                    */
/* ------------------------------------------------------------------------ */
#define RT(p, s) rand[(size_t)(p) + (size_t)(s) * (size_t)N]

void or_init(const or_params* pr, const float* rand, float* pos, float* vel,
             float* pbest, float* pbest_fit, float* gbest)
{
    const int N = pr->n_particles;
    for (int p = 0; p < N; ++p) {
        float P[7], V[7];
        int s = 0;
        for (int d = 0; d < 3; ++d) {
            P[d] = fmaf(pr->x_hi[d] - pr->x_lo[d], RT(p, s), pr->x_lo[d]); ++s;
            V[d] = fmaf(6.0f, RT(p, s), -3.0f); ++s; /* -vdot_ + 2 vdot_ r */
        }
        const float qinit[4] = {0.7071f, 0.7071f, 0.0f, 0.0f};
        float wp[3];
        for (int d = 0; d < 3; ++d) { wp[d] = fmaf(20.0f, RT(p, s), -10.0f); ++s; }
        float q[4];
        kinematics_from(qinit, q, wp);
        for (int i = 0; i < 4; ++i) P[3 + i] = q[i];
        V[3] = 0.0f;
        for (int d = 0; d < 3; ++d) { V[4 + d] = fmaf(20.0f, RT(p, s), -10.0f); ++s; }
        for (int d = 0; d < 7; ++d) {
            pos[p + d * N] = P[d];
            vel[p + d * N] = V[d];
            pbest[p + d * N] = P[d];
        }
        pbest_fit[p] = OR_BEST_FIT_INIT;
    }
    for (int d = 0; d < 7; ++d) gbest[d] = 0.0f;
    gbest[7] = OR_BEST_FIT_INIT;
}

/* UpdatePersonalAndGlobalBest (main.cu:768-855); gbest[8] = pose + fit. */
void or_update_best(int N, const float* fit, float* pbest_fit, const float* pos,
                    float* pbest, float* gbest, int best)
{
    for (int p = 0; p < N; ++p) {
        if (fit[p] < pbest_fit[p]) {
            pbest_fit[p] = fit[p];
            for (int d = 0; d < 7; ++d) pbest[p + d * N] = pos[p + d * N];
        }
    }
    if (fit[best] < gbest[7]) {
        gbest[7] = fit[best];
        for (int d = 0; d < 7; ++d) gbest[d] = pos[best + d * N];
    }
}

/* UpdateParticlesEquationKernel (main.cu:858-1060). */
void or_update_particles(const or_params* pr, const float* rand, int rand_idx,
                         float* pos, float* vel, const float* pbest, const float* gbest)
{
    const int N = pr->n_particles;
    for (int p = 0; p < N; ++p) {
        float P[7], V[7], B[7];
        for (int d = 0; d < 7; ++d) {
            P[d] = pos[p + d * N];
            V[d] = vel[p + d * N];
            B[d] = pbest[p + d * N];
        }
        int s = rand_idx;
        for (int d = 0; d < 3; ++d) {
            const float r1 = pr->c1 * RT(p, s); ++s;
            const float r2 = pr->c2 * RT(p, s); ++s;
            const float iV = (float)(pr->w * (double)V[d]);
            float v = fmaf(r1, B[d] - P[d], iV);
            v = fmaf(r2, gbest[d] - P[d], v);
            V[d] = v;
            P[d] += v;
            if (P[d] < pr->x_lo[d]) { P[d] = pr->x_lo[d]; V[d] = 0.0f; }
            else if (P[d] > pr->x_hi[d]) { P[d] = pr->x_hi[d]; V[d] = 0.0f; }
        }
        float wc[3] = {0, 0, 0}, ws[3] = {0, 0, 0};
        const float* targets[2] = {&B[3], &gbest[3]};
        float* outs[2] = {wc, ws};
        for (int k = 0; k < 2; ++k) {
            float qt[4];
            qprod_conj(targets[k], &P[3], qt);
            top_half(qt);
            float n2 = fmaf(qt[1], qt[1], qt[2] * qt[2]);
            n2 = fmaf(qt[3], qt[3], n2);
            const float n = sqrtf(n2);
            if (n > 1.0e-7f) {
                const float sc = (2.0f * or_acosf(qt[0])) * or_rcp(n);
                outs[k][0] = qt[1] * sc;
                outs[k][1] = qt[2] * sc;
                outs[k][2] = qt[3] * sc;
            }
        }
        V[3] = 0.0f;
        for (int i = 0; i < 3; ++i) {
            const float r1 = pr->c1 * RT(p, s); ++s;
            const float r2 = pr->c2 * RT(p, s); ++s;
            const float iW = (float)(pr->w * (double)V[4 + i]);
            float v = fmaf(r1, wc[i], iW);
            V[4 + i] = fmaf(r2, ws[i], v);
        }
        or_kinematics(&P[3], &V[4]);
        for (int d = 0; d < 7; ++d) {
            pos[p + d * N] = P[d];
            vel[p + d * N] = V[d];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* whole solve: host loop main.cu:1600-1733                                 */
/* ------------------------------------------------------------------------ */
int or_solve(const float* mesh, int F, const uint16_t* obs, const or_params* pr,
             const float* rand, int n_threads, or_solve_out* out)
{
    const int N = pr->n_particles, I = pr->n_iters;
    if (N <= 0 || I < 0 || F <= 0 || pr->num_points == 0) return -1;
    float* pos = (float*)malloc(sizeof(float) * 7 * (size_t)N);
    float* vel = (float*)malloc(sizeof(float) * 7 * (size_t)N);
    float* pbest = (float*)malloc(sizeof(float) * 7 * (size_t)N);
    float* pbest_fit = (float*)malloc(sizeof(float) * (size_t)N);
    float* fit = (float*)malloc(sizeof(float) * (size_t)N);
    uint64_t* rpx = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    uint64_t* vpx = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    int* degen = (int*)calloc((size_t)N, sizeof(int));
    int32_t* frames = (int32_t*)malloc(sizeof(int32_t) * OR_FRAME * (size_t)N);
    if (!pos || !vel || !pbest || !pbest_fit || !fit || !frames || !rpx || !vpx || !degen) return -2;
    for (size_t i = 0; i < OR_FRAME * (size_t)N; ++i) frames[i] = OR_EMPTY; /* main.cu:1588 */
    float gbest[8];
    or_init(pr, rand, pos, vel, pbest, pbest_fit, gbest);
    out->raster_px = out->visit_px = 0;
    out->n_degenerate = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    for (int it = 0; it < I; ++it) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
        for (int p = 0; p < N; ++p) {
            float pose[7];
            for (int d = 0; d < 7; ++d) pose[d] = pos[p + d * N];
            or_render_info ri;
            or_fitness_info fi;
            int32_t* fr = frames + (size_t)p * OR_FRAME;
            or_render_ex(mesh, F, pr->face_cap, pose, &pr->cam, pr->rect, pr->wrap, fr, &ri);
            fit[p] = or_fitness_ex(&ri, pr->rect, pr->num_points, pr->wrap, fr, obs, &fi);
            rpx[p] += ri.raster_px;
            vpx[p] += fi.visit_px;
            degen[p] += ri.degenerate;
        }
        if (out->fit_trace) memcpy(out->fit_trace + (size_t)it * N, fit, sizeof(float) * N);
        const int best = or_argmin(fit, N);
        if (out->argmin_trace) out->argmin_trace[it] = best;
        or_update_best(N, fit, pbest_fit, pos, pbest, gbest, best);
        if (out->gbest_trace) memcpy(out->gbest_trace + (size_t)it * 8, gbest, sizeof(float) * 8);
        or_update_particles(pr, rand, 12 + 12 * it, pos, vel, pbest, gbest);
    }
    for (int p = 0; p < N; ++p) {
        out->raster_px += rpx[p];
        out->visit_px += vpx[p];
        out->n_degenerate += degen[p];
    }
    memcpy(out->best_pose, gbest, sizeof(float) * 7);
    out->best_fit = gbest[7];
    if (out->pos) memcpy(out->pos, pos, sizeof(float) * 7 * N);
    if (out->vel) memcpy(out->vel, vel, sizeof(float) * 7 * N);
    if (out->pbest) memcpy(out->pbest, pbest, sizeof(float) * 7 * N);
    if (out->pbest_fit) memcpy(out->pbest_fit, pbest_fit, sizeof(float) * N);
    if (out->frames) memcpy(out->frames, frames, sizeof(int32_t) * OR_FRAME * N);
    free(pos); free(vel); free(pbest); free(pbest_fit); free(fit);
    free(rpx); free(vpx); free(degen); free(frames);
    return 0;
}
