// qpso_math.h -- gfx950 device math of the Q-PSO hot path.
//
// Arithmetic contract (the kernels are compiled with -ffp-contract=off, so
// every fused multiply-add below is written out): each a*b+c site of the
// reference is one fmaf/fma, fusing the first product of a sum of two
// products; __frcp_rn is a correctly rounded 1/x; sqrtf is IEEE; sin, cos
// and acos are evaluated in binary64 from Taylor series after Cody-Waite
// reduction and rounded once to binary32.  Signed 32-bit integer arithmetic
// wraps (done on uint32_t).
#ifndef QPSO_MATH_H_
#define QPSO_MATH_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qpso {

#define QD __device__ __forceinline__

QD int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
QD int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
QD int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
QD int32_t wshl(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }

// __float2int_rn: round half to even, saturating, NaN -> 0 (main.h:1331-1338).
// v_rndne_f32 then v_cvt_i32_f32: the CDNA conversion clamps out-of-range
// values to INT32_MIN/INT32_MAX and converts NaN to 0, which is exactly the
// saturating semantics above (a C cast would be undefined out of range, so the
// conversion is issued directly; it is a plain VALU op).
QD int32_t rint_sat(float x) {
    const float r = __builtin_rintf(x);
    int32_t v;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(v) : "v"(r));
    return v;
}

QD float rcp(float x) { return 1.0f / x; }  // __frcp_rn

// ---- sin / cos / acos (binary64 evaluation, one final rounding) ----------
constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
constexpr double kPio2Lo = 0x1.1a62633145c07p-54;
constexpr double kPi = 0x1.921fb54442d18p+1;
constexpr double k2OverPi = 0x1.45f306dc9c883p-1;

QD double sin_poly(double r) {
    // (-1)^n / (2n+1)!, n = 1..7
    const double z = r * r;
    double p = -0x1.ae7f3e733b81fp-41;
    p = __builtin_fma(p, z, 0x1.6124613a86d09p-33);
    p = __builtin_fma(p, z, -0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, z, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, z, -0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, z, 0x1.1111111111111p-7);
    p = __builtin_fma(p, z, -0x1.5555555555555p-3);
    return __builtin_fma(r * z, p, r);
}

QD double cos_poly(double r) {
    // (-1)^n / (2n)!, n = 1..8
    const double z = r * r;
    double p = 0x1.ae7f3e733b81fp-45;
    p = __builtin_fma(p, z, -0x1.93974a8c07c9dp-37);
    p = __builtin_fma(p, z, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, z, -0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, z, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, z, -0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, z, 0x1.5555555555555p-5);
    p = __builtin_fma(p, z, -0x1.0000000000000p-1);
    return __builtin_fma(z, p, 1.0);
}

QD double reduce_pio2(double x, int& q) {
    const double k = __builtin_rint(x * k2OverPi);
    double r = __builtin_fma(-k, kPio2Hi, x);
    r = __builtin_fma(-k, kPio2Lo, r);
    q = (int)(k - 4.0 * __builtin_floor(k * 0.25));
    return r;
}

QD float sinf_q(float xf) {
    const double x = (double)xf;
    if (!__builtin_isfinite(x)) return (float)(x - x);
    int q;
    const double r = reduce_pio2(x, q);
    const double s = sin_poly(r), c = cos_poly(r);
    const double v = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
    return (float)v;
}

QD float cosf_q(float xf) {
    const double x = (double)xf;
    if (!__builtin_isfinite(x)) return (float)(x - x);
    int q;
    const double r = reduce_pio2(x, q);
    const double s = sin_poly(r), c = cos_poly(r);
    const double v = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
    return (float)v;
}

QD double asin_core(double s, double z) {
    // (2n)! / (4^n (n!)^2 (2n+1)), n = 1..20, Horner in z = s^2
    double p = 0x1.90cb77f60c7cep-9;
    p = __builtin_fma(p, z, 0x1.b026f57b13b14p-9);
    p = __builtin_fma(p, z, 0x1.d3d2a8e0dd67dp-9);
    p = __builtin_fma(p, z, 0x1.fcaf8fb6db6dbp-9);
    p = __builtin_fma(p, z, 0x1.15ee9d45d1746p-8);
    p = __builtin_fma(p, z, 0x1.31683bdef7bdfp-8);
    p = __builtin_fma(p, z, 0x1.51ba308d3dcb1p-8);
    p = __builtin_fma(p, z, 0x1.782dda12f684cp-8);
    p = __builtin_fma(p, z, 0x1.a6863d70a3d71p-8);
    p = __builtin_fma(p, z, 0x1.df3bd37a6f4dfp-8);
    p = __builtin_fma(p, z, 0x1.12ef3cf3cf3cfp-7);
    p = __builtin_fma(p, z, 0x1.3fde50d79435ep-7);
    p = __builtin_fma(p, z, 0x1.7a87878787878p-7);
    p = __builtin_fma(p, z, 0x1.c99999999999ap-7);
    p = __builtin_fma(p, z, 0x1.1c4ec4ec4ec4fp-6);
    p = __builtin_fma(p, z, 0x1.6e8ba2e8ba2e9p-6);
    p = __builtin_fma(p, z, 0x1.f1c71c71c71c7p-6);
    p = __builtin_fma(p, z, 0x1.6db6db6db6db7p-5);
    p = __builtin_fma(p, z, 0x1.3333333333333p-4);
    p = __builtin_fma(p, z, 0x1.5555555555555p-3);
    return __builtin_fma(s * z, p, s);
}

QD float acosf_q(float xf) {
    const double x = (double)xf;
    if (x != x) return xf;
    const double ax = __builtin_fabs(x);
    if (ax > 1.0) return __builtin_nanf("");
    if (ax <= 0.5) return (float)(kPio2Hi - asin_core(x, x * x));
    const double z = (1.0 - ax) * 0.5;
    const double s = __builtin_sqrt(z);
    const double a2 = 2.0 * asin_core(s, z);
    return (float)(x > 0.0 ? a2 : kPi - a2);
}

// ---- quaternions: [q0, q1, q2, q3], scalar first (main.h:208-413) --------
// QuatProd_pqPSOstateVect (main.h:249-255)
QD void qprod(const float p[4], const float q[4], float o[4]) {
    float a;
    a = __builtin_fmaf(p[0], q[0], -(p[1] * q[1])); a = __builtin_fmaf(-p[2], q[2], a); o[0] = __builtin_fmaf(-p[3], q[3], a);
    a = __builtin_fmaf(p[1], q[0], p[0] * q[1]);    a = __builtin_fmaf(p[2], q[3], a);  o[1] = __builtin_fmaf(-p[3], q[2], a);
    a = __builtin_fmaf(p[2], q[0], p[0] * q[2]);    a = __builtin_fmaf(p[3], q[1], a);  o[2] = __builtin_fmaf(-p[1], q[3], a);
    a = __builtin_fmaf(p[3], q[0], p[0] * q[3]);    a = __builtin_fmaf(p[1], q[2], a);  o[3] = __builtin_fmaf(-p[2], q[1], a);
}

// QuatProd_pq_1PSOstateVect (main.h:257-264): p (x) conj(q)
QD void qprod_conj(const float p[4], const float q[4], float o[4]) {
    float a;
    a = __builtin_fmaf(p[0], q[0], p[1] * q[1]);    a = __builtin_fmaf(p[2], q[2], a);  o[0] = __builtin_fmaf(p[3], q[3], a);
    a = __builtin_fmaf(p[1], q[0], -(p[0] * q[1])); a = __builtin_fmaf(-p[2], q[3], a); o[1] = __builtin_fmaf(p[3], q[2], a);
    a = __builtin_fmaf(p[2], q[0], -(p[0] * q[2])); a = __builtin_fmaf(-p[3], q[1], a); o[2] = __builtin_fmaf(p[1], q[3], a);
    a = __builtin_fmaf(p[3], q[0], -(p[0] * q[3])); a = __builtin_fmaf(-p[1], q[2], a); o[3] = __builtin_fmaf(p[2], q[1], a);
}

QD void top_half(float q[4]) {  // Cast2TopHalfHypersphere (main.h:224-246)
    if (q[0] < 0.0f) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; q[3] = -q[3]; }
}

QD void normalize4(float q[4]) {  // NormalizeQuaternionPSOstateVect (main.h:210-221)
    float s = __builtin_fmaf(q[0], q[0], q[1] * q[1]);
    s = __builtin_fmaf(q[2], q[2], s);
    s = __builtin_fmaf(q[3], q[3], s);
    const float d = __builtin_sqrtf(s);
    if (d > 1.0e-7f) {
        const float inv = rcp(d);
        q[0] *= inv; q[1] *= inv; q[2] *= inv; q[3] *= inv;
    }
}

// QuatKinematicsPSOStateVect / ...WithInit (main.h:267-307, 374-413)
QD void kinematics_from(const float src[4], float dst[4], const float w[3]) {
    const float wt[4] = {0.0f, w[0] * 0.5f, w[1] * 0.5f, w[2] * 0.5f};
    float n2 = __builtin_fmaf(wt[1], wt[1], wt[2] * wt[2]);
    n2 = __builtin_fmaf(wt[3], wt[3], n2);
    const float d = __builtin_sqrtf(n2) * 0.5f;
    const float cw = cosf_q(d);
    const float sw = (d < 1.0e-7f) ? 1.0f : sinf_q(d) * rcp(d);
    float qxw[4];
    qprod(src, wt, qxw);
    const float hs = 0.5f * sw;
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[i] = __builtin_fmaf(cw, src[i], hs * qxw[i]);
    normalize4(dst);
    top_half(dst);
}

// ---- projection (main.h:514-545 with main.h:309-369) ----------------------
struct Cam {
    double fx, fy, cx, cy;
};

QD void project(const float pose[7], float vx, float vy, float vz, const Cam& c,
                float& u, float& v, float& Z) {
    const float q0 = pose[3], q1 = pose[4], q2 = pose[5], q3 = pose[6];
    const float t0 = 2.0f * __builtin_fmaf(q2, vz, -(q3 * vy));
    const float t1 = 2.0f * -__builtin_fmaf(q1, vz, -(q3 * vx));
    const float t2 = 2.0f * __builtin_fmaf(q1, vy, -(q2 * vx));
    const float k0 = __builtin_fmaf(q2, t2, -(q3 * t1));
    const float k1 = -__builtin_fmaf(q1, t2, -(q3 * t0));
    const float k2 = __builtin_fmaf(q1, t1, -(q2 * t0));
    float px = __builtin_fmaf(q0, t0, vx) + k0;
    float py = __builtin_fmaf(q0, t1, vy) + k1;
    float pz = __builtin_fmaf(q0, t2, vz) + k2;
    px += pose[0];
    py += pose[1];
    pz += pose[2];
    const float r = rcp(pz);
    u = (float)__builtin_fma(c.fx, (double)(px * r), c.cx);
    v = (float)__builtin_fma(c.fy, (double)(py * r), c.cy);
    Z = 1000.0f * pz;
}

// Fixed-point face: 28.4 vertices and the integer bbox (main.h:1326-1353).
struct Face {
    int32_t X[3], Y[3];
    float Z[3];
    int32_t minx, maxx, miny, maxy;
    bool degen;
};

QD void face_bbox(Face& fc) {
    fc.minx = wadd(min(min(fc.X[0], fc.X[1]), fc.X[2]), 15) >> 4;
    fc.maxx = wadd(max(max(fc.X[0], fc.X[1]), fc.X[2]), 15) >> 4;
    fc.miny = wadd(min(min(fc.Y[0], fc.Y[1]), fc.Y[2]), 15) >> 4;
    fc.maxy = wadd(max(max(fc.Y[0], fc.Y[1]), fc.Y[2]), 15) >> 4;
}

// Projected vertex in 28.4 fixed point; `degen` if outside the +-2^30 range.
QD bool vertex_fixed(const float pose[7], float vx, float vy, float vz, const Cam& c,
                     int32_t& X, int32_t& Y, float& Z) {
    float u, v;
    project(pose, vx, vy, vz, c, u, v, Z);
    X = rint_sat(16.0f * u);
    Y = rint_sat(16.0f * v);
    return (X > QPSO_FXP_LIMIT) | (X < -QPSO_FXP_LIMIT) | (Y > QPSO_FXP_LIMIT) | (Y < -QPSO_FXP_LIMIT);
}

// Face f projected from the corner table (corner j of face f at
// corners[12 f + 4 j], (x, y, z, vertex index): one 16-byte load per corner).
QD void face_project(const float* __restrict__ corners, int f, const float pose[7], const Cam& c, Face& fc) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4* __restrict__ cr = reinterpret_cast<const f4*>(corners) + 3 * (size_t)f;
    fc.degen = false;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const f4 w = cr[j];
        float u, v, Z;
        project(pose, w.x, w.y, w.z, c, u, v, Z);
        fc.X[j] = rint_sat(16.0f * u);
        fc.Y[j] = rint_sat(16.0f * v);
        fc.Z[j] = Z;
        fc.degen |= (fc.X[j] > QPSO_FXP_LIMIT) | (fc.X[j] < -QPSO_FXP_LIMIT) |
                    (fc.Y[j] > QPSO_FXP_LIMIT) | (fc.Y[j] < -QPSO_FXP_LIMIT);
    }
    face_bbox(fc);
}

// Raster set-up (main.h:1355-1419): edge functions E23 and E31 at the bbox
// origin and their per-pixel steps; E12 never influences the output because
// the inside test of the reference is always true (bbox fill).
struct Setup {
    int32_t CY2, CY3, FDX23, FDX31, FDY23, FDY31;
    float z1, z2, z3;
};

// The faces face_setup accepts: positive signed area (counter-clockwise in
// 28.4 screen coordinates) and a non-empty bbox.
QD bool face_front(const Face& fc) {
    const int32_t DX12 = wsub(fc.X[0], fc.X[1]), DX31 = wsub(fc.X[2], fc.X[0]);
    const int32_t DY12 = wsub(fc.Y[0], fc.Y[1]), DY31 = wsub(fc.Y[2], fc.Y[0]);
    return wsub(wmul(DX12, DY31), wmul(DX31, DY12)) > 0 && fc.maxx > fc.minx && fc.maxy > fc.miny;
}

QD bool face_setup(const Face& fc, Setup& s) {
    const int32_t X1 = fc.X[0], X2 = fc.X[1], X3 = fc.X[2];
    const int32_t Y1 = fc.Y[0], Y2 = fc.Y[1], Y3 = fc.Y[2];
    const int32_t DX12 = wsub(X1, X2), DX23 = wsub(X2, X3), DX31 = wsub(X3, X1);
    const int32_t DY12 = wsub(Y1, Y2), DY23 = wsub(Y2, Y3), DY31 = wsub(Y3, Y1);
    const int32_t tri = wsub(wmul(DX12, DY31), wmul(DX31, DY12));
    if (tri <= 0) return false;
    if (fc.maxx - fc.minx <= 0 || fc.maxy - fc.miny <= 0) return false;
    const float inv = 1.0f * rcp((float)tri);
    s.FDX23 = wshl(DX23, 4); s.FDX31 = wshl(DX31, 4);
    s.FDY23 = wshl(DY23, 4); s.FDY31 = wshl(DY31, 4);
    int32_t C2 = wsub(wmul(DY23, X2), wmul(DX23, Y2));
    int32_t C3 = wsub(wmul(DY31, X3), wmul(DX31, Y3));
    C2 = wadd(C2, (DY23 < 0 || (DY23 == 0 && DX23 > 0)) ? 1 : 0);
    C3 = wadd(C3, (DY31 < 0 || (DY31 == 0 && DX31 > 0)) ? 1 : 0);
    const int32_t MINY = wshl(fc.miny, 4), MINX = wshl(fc.minx, 4);
    s.CY2 = wsub(wadd(C2, wmul(DX23, MINY)), wmul(DY23, MINX));
    s.CY3 = wsub(wadd(C3, wmul(DX31, MINY)), wmul(DY31, MINX));
    s.z1 = fc.Z[0];
    s.z2 = (fc.Z[1] - fc.Z[0]) * inv;
    s.z3 = (fc.Z[2] - fc.Z[0]) * inv;
    return true;
}

// Depth written at a pixel with edge values CX2, CX3 (main.h:1433-1447).
QD int32_t pixel_depth(int32_t CX2, int32_t CX3, const Setup& s) {
    const int32_t cx2 = wadd(CX2, 15) >> 8;
    const int32_t cx3 = wadd(CX3, 15) >> 8;
    float d = __builtin_fmaf((float)cx2, s.z2, s.z1);
    d = __builtin_fmaf((float)cx3, s.z3, d);
    return rint_sat(d);
}

// Range [lo, hi] of cx = (CX + 15) >> 8 over the offsets [ia, ib] x [ja, jb]
// of an edge CX = CY + FDX j - FDY i, in 64 bits; false if a corner leaves
// [INT32_MIN, INT32_MAX - 15], where the 32-bit value would wrap.
QD bool edge_range64(int32_t CY, int32_t FDX, int32_t FDY, int32_t ia, int32_t ib, int32_t ja, int32_t jb,
                     int32_t& lo, int32_t& hi) {
    const long long e = (long long)CY + (long long)FDX * ja - (long long)FDY * ia;
    const long long dx = -(long long)FDY * (ib - ia), dy = (long long)FDX * (jb - ja);
    const long long mn = e + (dx < 0 ? dx : 0) + (dy < 0 ? dy : 0);
    const long long mx = e + (dx > 0 ? dx : 0) + (dy > 0 ? dy : 0);
    if (mn < (long long)INT32_MIN || mx > (long long)INT32_MAX - 15) return false;
    lo = (int32_t)((mn + 15) >> 8);
    hi = (int32_t)((mx + 15) >> 8);
    return true;
}

// The same in 32 bits, for |FDX|, |FDY| < 2^18, |CY| < 2^28 and offsets below
// 1024: every product is below 2^28 in magnitude and every corner value below
// 5 * 2^28, so nothing wraps; the operands fit the 24-bit multiply.
QD void edge_range32(int32_t CY, int32_t FDX, int32_t FDY, int32_t ia, int32_t ib, int32_t ja, int32_t jb,
                     int32_t& lo, int32_t& hi) {
    const int32_t e = CY + __mul24(FDX, ja) - __mul24(FDY, ia);
    const int32_t dx = -__mul24(FDY, ib - ia), dy = __mul24(FDX, jb - ja);
    lo = (e + min(dx, 0) + min(dy, 0) + 15) >> 8;
    hi = (e + max(dx, 0) + max(dy, 0) + 15) >> 8;
}

// Exact bounds [kmin, kmax] of the depth a face writes over the pixel
// rectangle [x0, x1) x [y0, y1) of its bbox (false where they cannot be had
// this way: a wrapping edge value or a NaN corner).  The edge value at a pixel is
// CY + FDX (y - miny) - FDY (x - minx) mod 2^32; if that affine function stays
// inside [INT32_MIN, INT32_MAX - 15] at the four corners it does so everywhere,
// no wrap occurs, and cx = (CX + 15) >> 8 ranges between its corner values.
// The depth fma(cx3, z3, fma(cx2, z2, z1)) is monotone in cx2, cx3 (for fixed
// z2, z3) and in the addend, and rint is monotone, so the corner combinations
// bound every pixel's written value exactly.  The rectangle lies inside the
// bbox, so the offsets are >= 0.
QD bool depth_bounds(const Face& fc, const Setup& s, int32_t x0, int32_t x1, int32_t y0, int32_t y1,
                     int32_t& kmin, int32_t& kmax) {
    const int32_t ia = x0 - fc.minx, ib = x1 - 1 - fc.minx, ja = y0 - fc.miny, jb = y1 - 1 - fc.miny;
    const uint32_t steps = (uint32_t)__builtin_abs(s.FDX23) | (uint32_t)__builtin_abs(s.FDY23) |
                           (uint32_t)__builtin_abs(s.FDX31) | (uint32_t)__builtin_abs(s.FDY31);
    const uint32_t origin = (uint32_t)__builtin_abs(s.CY2) | (uint32_t)__builtin_abs(s.CY3);
    int32_t lo2, hi2, lo3, hi3;
    if (steps < (1u << 18) && origin < (1u << 28) && (uint32_t)(ib | jb) < 1024u) {
        edge_range32(s.CY2, s.FDX23, s.FDY23, ia, ib, ja, jb, lo2, hi2);
        edge_range32(s.CY3, s.FDX31, s.FDY31, ia, ib, ja, jb, lo3, hi3);
    } else if (!edge_range64(s.CY2, s.FDX23, s.FDY23, ia, ib, ja, jb, lo2, hi2) ||
               !edge_range64(s.CY3, s.FDX31, s.FDY31, ia, ib, ja, jb, lo3, hi3)) {
        return false;
    }
    const float fa = __builtin_fmaf((float)lo2, s.z2, s.z1), fb = __builtin_fmaf((float)hi2, s.z2, s.z1);
    const float d0 = __builtin_fmaf((float)lo3, s.z3, fa), d1 = __builtin_fmaf((float)lo3, s.z3, fb);
    const float d2 = __builtin_fmaf((float)hi3, s.z3, fa), d3 = __builtin_fmaf((float)hi3, s.z3, fb);
    if (d0 != d0 || d1 != d1 || d2 != d2 || d3 != d3) return false;
    kmin = rint_sat(__builtin_fminf(__builtin_fminf(d0, d1), __builtin_fminf(d2, d3)));
    kmax = rint_sat(__builtin_fmaxf(__builtin_fmaxf(d0, d1), __builtin_fmaxf(d2, d3)));
    return true;
}

// Written depth at pixel (x, y) of the face's bbox, from the edge values there.
QD int32_t depth_at(const Face& fc, const Setup& s, int32_t x, int32_t y) {
    const int32_t CX2 = wsub(wadd(s.CY2, wmul(y - fc.miny, s.FDX23)), wmul(x - fc.minx, s.FDY23));
    const int32_t CX3 = wsub(wadd(s.CY3, wmul(y - fc.miny, s.FDX31)), wmul(x - fc.minx, s.FDY31));
    return pixel_depth(CX2, CX3, s);
}

// Fitness row index y(t) = trunc(f32(t) * rcp(f32(W))) (main.cu:645).
QD int32_t row_of(int32_t t, float rW) { return (int32_t)((float)t * rW); }  // the conversion truncates

#undef QD
}  // namespace qpso

#endif
