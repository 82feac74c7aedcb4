// Scalar and vector arithmetic of the step core, written once for the host and for gfx950.
//
// Every function here is evaluated in IEEE single precision with the exact operation order the reference uses
// (no contraction: the build passes -ffp-contract=off; division and sqrt are correctly rounded on both sides).
// The transcendental kernels restate the reference's own sequences (include/rocketgpu/detail/math.hpp:17-492): the
// Windows UCRT float kernels the game links, evaluated in double with explicit fused multiply-adds.  Their
// polynomial coefficients are the UCRT's; the lookup tables are generated from their mathematical definition by
// gen_math_tables.py.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "rocketgpu_amd/rg_types.h"

namespace rg {
namespace core {

// ------------------------------------------------------------------------------------------------------------
// Scalars.  Min/Max/Abs keep the reference's comparison forms (common.hpp:160-174), which fix the result for
// signed zeros and NaNs.

RG_HD constexpr float Abs(float v) { return v >= 0.0f ? v : -v; }
RG_HD constexpr float Min(float a, float b) { return a < b ? a : b; }
RG_HD constexpr float Max(float a, float b) { return a > b ? a : b; }
RG_HD constexpr int MinI(int a, int b) { return a < b ? a : b; }
RG_HD constexpr int MaxI(int a, int b) { return a > b ? a : b; }
RG_HD constexpr float Clamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
RG_HD constexpr float Pi() { return 3.14159265358979323846f; }
RG_HD constexpr float HalfPi() { return 1.57079632679489661923f; }
constexpr float kSimdEpsilon = 1.1920928955078125e-7f;  // FLT_EPSILON, Bullet's SIMD_EPSILON
constexpr float kUuToBt = 1.0f / 50.0f;
constexpr float kBtToUu = 50.0f;

RG_HD uint32_t FloatBits(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof(u));
    return u;
}
RG_HD float BitsFloat(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
}
RG_HD uint64_t DoubleBits(double d) {
    uint64_t u;
    memcpy(&u, &d, sizeof(u));
    return u;
}
RG_HD double BitsDouble(uint64_t u) {
    double d;
    memcpy(&d, &u, sizeof(d));
    return d;
}

RG_HD int Ctz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ctz(x);
#else
    return __builtin_ctz(x);
#endif
}

RG_HD int Ctz64(uint64_t x) {
    const uint32_t lo = static_cast<uint32_t>(x);
    return lo != 0u ? Ctz32(lo) : 32 + Ctz32(static_cast<uint32_t>(x >> 32));
}

// ------------------------------------------------------------------------------------------------------------
// Vec3 / Quat.

RG_HD constexpr Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
RG_HD constexpr Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
RG_HD constexpr Vec3 operator-(const Vec3& v) { return Vec3(-v.x, -v.y, -v.z); }
RG_HD constexpr Vec3 operator*(const Vec3& a, float s) { return Vec3(a.x * s, a.y * s, a.z * s); }
RG_HD constexpr Vec3 operator*(float s, const Vec3& a) { return a * s; }
RG_HD constexpr Vec3 operator/(const Vec3& a, float s) { return Vec3(a.x / s, a.y / s, a.z / s); }
RG_HD Vec3& operator+=(Vec3& a, const Vec3& b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
RG_HD Vec3& operator-=(Vec3& a, const Vec3& b) {
    a.x -= b.x;
    a.y -= b.y;
    a.z -= b.z;
    return a;
}
RG_HD constexpr float Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RG_HD constexpr Vec3 Cross(const Vec3& a, const Vec3& b) {
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RG_HD constexpr Vec3 Mul3(const Vec3& a, const Vec3& b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
RG_HD constexpr Vec3 Div3(const Vec3& a, const Vec3& b) { return Vec3(a.x / b.x, a.y / b.y, a.z / b.z); }
RG_HD constexpr Vec3 Min3(const Vec3& a, const Vec3& b) { return Vec3(Min(a.x, b.x), Min(a.y, b.y), Min(a.z, b.z)); }
RG_HD constexpr Vec3 Max3(const Vec3& a, const Vec3& b) { return Vec3(Max(a.x, b.x), Max(a.y, b.y), Max(a.z, b.z)); }
RG_HD float LengthSq(const Vec3& v) { return Dot(v, v); }
RG_HD float Length(const Vec3& v) { return sqrtf(LengthSq(v)); }
RG_HD float Rsqrt(float v) { return 1.0f / sqrtf(v); }
RG_HD Vec3 Normalize(const Vec3& v, const Vec3& fallback = Vec3(0.0f, 0.0f, 0.0f)) {
    const float len_sq = LengthSq(v);
    if (len_sq <= 1e-12f) return fallback;
    return v * Rsqrt(len_sq);
}
// The game's replicated-state quantization (raycast_wheels_and_contact_prep.hpp:3862-3875): positions to 0.01 uu
// rounded, velocities truncated with a tenth-step offset away from zero; and a length cap.
RG_HD float QuantRoundPos(float x, float scale) { return floorf(x * scale + 0.5f) * (1.0f / scale); }

RG_HD float QuantTruncOffset(float x, float scale) {
    const float inv_scale = 1.0f / scale;
    const float k = static_cast<float>(static_cast<int32_t>(x * scale));
    const float rounded = k * inv_scale;
    const float mag = 0.1f * inv_scale;
    const float sgn = rounded == 0.0f ? 0.0f : (rounded > 0.0f ? 1.0f : -1.0f);
    return rounded + sgn * mag;
}

RG_HD Vec3 ClampLength(const Vec3& v, float max_len) {
    const float len_sq = LengthSq(v);
    if (len_sq <= max_len * max_len) return v;
    return (v * Rsqrt(len_sq)) * max_len;
}

RG_HD Vec3 ComponentsBtFromUu(const Vec3& v) { return Vec3(v.x / 50.0f, v.y / 50.0f, v.z / 50.0f); }
RG_HD bool SameBits(float a, float b) { return FloatBits(a) == FloatBits(b); }
RG_HD bool SameBits(const Vec3& a, const Vec3& b) {
    return SameBits(a.x, b.x) && SameBits(a.y, b.y) && SameBits(a.z, b.z);
}

RG_HD float QuatLengthSqBullet(const Quat& q) { return (q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w); }

// Rotation-matrix columns of q in Bullet's btMatrix3x3::setRotation form (math.hpp:956-982).
RG_HD void QuatColumns(const Quat& q, Vec3& fwd, Vec3& right, Vec3& up) {
    const float len_sq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    if (len_sq <= 1e-12f) {
        fwd = Vec3(1.0f, 0.0f, 0.0f);
        right = Vec3(0.0f, 1.0f, 0.0f);
        up = Vec3(0.0f, 0.0f, 1.0f);
        return;
    }
    const float s = 2.0f / len_sq;
    const float xs = q.x * s, ys = q.y * s, zs = q.z * s;
    const float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
    const float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
    const float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
    fwd = Vec3(1.0f - (yy + zz), xy + wz, xz - wy);
    right = Vec3(xy - wz, 1.0f - (xx + zz), yz + wx);
    up = Vec3(xz + wy, yz - wx, 1.0f - (xx + yy));
}

// The game's rotation-matrix -> quaternion (math.hpp:1092-1119): Shepperd's method with the game's branch order.
RG_HD Quat GameMatrixToQuat(const Vec3& fwd, const Vec3& right, const Vec3& up) {
    const float m00 = fwd.x, m01 = right.x, m02 = up.x;
    const float m10 = fwd.y, m11 = right.y, m12 = up.y;
    const float m20 = fwd.z, m21 = right.z, m22 = up.z;
    const float trace = (m00 + m11) + m22;
    if (trace > 0.0f) {
        const float r = sqrtf(trace + 1.0f);
        const float inv = 0.5f / r;
        return Quat(r * 0.5f, (m21 - m12) * inv, (m02 - m20) * inv, (m10 - m01) * inv);
    }
    if (m00 < m11 ? m11 < m22 : m00 < m22) {
        const float r = sqrtf(((m22 - m00) - m11) + 1.0f);
        const float inv = 0.5f / r;
        return Quat((m10 - m01) * inv, (m02 + m20) * inv, (m12 + m21) * inv, r * 0.5f);
    }
    if (m00 < m11) {
        const float r = sqrtf(((m11 - m22) - m00) + 1.0f);
        const float inv = 0.5f / r;
        return Quat((m02 - m20) * inv, (m01 + m10) * inv, r * 0.5f, (m21 + m12) * inv);
    }
    const float r = sqrtf(((m00 - m11) - m22) + 1.0f);
    const float inv = 0.5f / r;
    return Quat((m21 - m12) * inv, r * 0.5f, (m10 + m01) * inv, (m20 + m02) * inv);
}

// The game's quaternion -> matrix columns (math.hpp:1121-1131).
RG_HD void GameQuatToColumns(const Quat& q, Vec3& fwd, Vec3& right, Vec3& up) {
    const float d = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    const float s = 2.0f / d;
    const float xs = q.x * s, ys = q.y * s, zs = q.z * s;
    const float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
    const float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
    const float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
    fwd = Vec3(1.0f - (yy + zz), xy + wz, xz - wy);
    right = Vec3(xy - wz, 1.0f - (xx + zz), yz + wx);
    up = Vec3(xz + wy, yz - wx, 1.0f - (xx + yy));
}

RG_HD Quat GameQuatNormalize(const Quat& q, float min_len_sq) {
    const float len_sq = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    if (!(len_sq > min_len_sq)) return Quat(1.0f, 0.0f, 0.0f, 0.0f);
    const float s = 1.0f / sqrtf(len_sq);
    return Quat(q.w * s, q.x * s, q.y * s, q.z * s);
}

// Smallest-three quaternion compression of the game's replicated rigid-body state (math.hpp:1160-1200): the three
// smaller components are sent as 18-bit fixed point over [-1/sqrt2, 1/sqrt2] and the largest is rebuilt.
RG_HD float GameQuatQuantizeComponent(float c) {
    constexpr int kMax = (1 << 18) - 1;
    const float max_f = static_cast<float>(kMax);
    float v = ((c * 1.41421354f) * 0.5f + 0.5f) * max_f;
    v = (v + v) + 0.5f;
    int k = static_cast<int>(rintf(v));
    k = k < 0 ? (k - 1) / 2 : k / 2;
    k = k < 0 ? 0 : (k < kMax ? k : kMax);
    const float r = static_cast<float>(k) / max_f - 0.5f;
    return (r + r) * 0.707106769f;
}

RG_HD Quat GameQuantizeQuat(const Quat& q_in) {
    const Quat q = GameQuatNormalize(q_in, 1e-8f);
    const float a0 = fabsf(q.x), a1 = fabsf(q.y), a2 = fabsf(q.z), a3 = fabsf(q.w);
    const int i01 = a0 < a1 ? 1 : 0;
    const int i23 = a2 < a3 ? 3 : 2;
    const int big = (i01 == 1 ? a1 : a0) >= (i23 == 3 ? a3 : a2) ? i01 : i23;
    const float c_big = big == 0 ? q.x : big == 1 ? q.y : big == 2 ? q.z : q.w;
    const float sign = c_big < 0.0f ? -1.0f : 1.0f;
    const float d0 = GameQuatQuantizeComponent((big == 0 ? q.y : q.x) * sign);
    const float d1 = GameQuatQuantizeComponent((big <= 1 ? q.z : q.y) * sign);
    const float d2 = GameQuatQuantizeComponent((big <= 2 ? q.w : q.z) * sign);
    const float dropped = sqrtf(((1.0f - d0 * d0) - d1 * d1) - d2 * d2);
    const float ox = big == 0 ? dropped : d0;
    const float oy = big == 1 ? dropped : (big == 0 ? d0 : d1);
    const float oz = big == 2 ? dropped : (big <= 1 ? d1 : d2);
    const float ow = big == 3 ? dropped : d2;
    return GameQuatNormalize(Quat(ow, ox, oy, oz), 1e-8f);
}

// Quantize a basis through the compressed quaternion and rebuild the columns from it.
RG_HD Quat GameQuantizeColumns(Vec3& fwd, Vec3& right, Vec3& up) {
    const Quat q = GameQuantizeQuat(GameMatrixToQuat(fwd, right, up));
    GameQuatToColumns(q, fwd, right, up);
    return q;
}

// ------------------------------------------------------------------------------------------------------------
// Transcendentals (UCRT float kernels in double arithmetic with explicit FMA).

// The tables are constexpr members so that host and gfx950 code read the same bits (clang places constexpr
// arrays ODR-used from device code in device memory).
struct MathTables {
#define RG_MATH_TABLE(type, name, n) static constexpr type name[n]
#include "rg_math_tables.inc"
#undef RG_MATH_TABLE
};

RG_HD double UcrtSinPoly(double r) {
    const double r2 = r * r;
    double p = fma(r2, 0x1.71de3a556c734p-19, -0x1.a01a01a01a01ap-13);
    p = fma(p, r2, 0x1.1111111111111p-7);
    p = fma(p, r2, -0x1.5555555555555p-3);
    return fma(p, r * r2, r);
}

RG_HD double UcrtCosPoly(double r) {
    const double r2 = r * r;
    const double base = 1.0 - r2 * 0.5;
    double p = fma(r2, -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a019p-16);
    p = fma(p, r2, -0x1.6c16c16c16c16p-10);
    p = fma(p, r2, 0x1.5555555555555p-5);
    return fma(p, r2 * r2, base);
}

// Cody-Waite reduction by pi/2 in two parts; quadrant in q (0..3).
RG_HD double UcrtReduceHalfPi(double a, int* q) {
    const int n = static_cast<int>(fma(a, 0x1.45f306dc9c883p-1, 0.5));
    const double nd = static_cast<double>(n);
    *q = n & 3;
    return fma(-nd, 0x1.921fb54400000p+0, a) - nd * 0x1.0b4611a626331p-34;
}

constexpr uint32_t kUcrtSinfLimitBits = 0x4b800456u;  // |x| at and above this use the large-argument path
constexpr uint32_t kUcrtCosfLimitBits = 0x4f490fdbu;

RG_HD float UcrtSinf(float x) {
    const uint32_t ax = FloatBits(x) & 0x7fffffffu;
    if (ax < 0x39000000u) return x;
    const double d = static_cast<double>(x);
    if (ax <= 0x3f490fdbu) {
        if (ax < 0x3c000000u) return static_cast<float>(fma(-((d * d) * d), 0x1.5555555555555p-3, d));
        return static_cast<float>(UcrtSinPoly(d));
    }
    int q = 0;
    const double r = UcrtReduceHalfPi(fabs(d), &q);
    const double y = (q & 1) ? UcrtCosPoly(r) : UcrtSinPoly(r);
    return static_cast<float>(((q >= 2) != (x < 0.0f)) ? -y : y);
}

RG_HD float UcrtCosf(float x) {
    const uint32_t ax = FloatBits(x) & 0x7fffffffu;
    if (ax < 0x39000000u) return 1.0f;
    const double d = static_cast<double>(x);
    if (ax <= 0x3f490fdbu) {
        if (ax < 0x3c000000u) return static_cast<float>(fma(-(d * 0.5), d, 1.0));
        return static_cast<float>(UcrtCosPoly(d));
    }
    int q = 0;
    const double r = UcrtReduceHalfPi(fabs(d), &q);
    const double y = (q & 1) ? UcrtSinPoly(r) : UcrtCosPoly(r);
    return static_cast<float>((q == 1 || q == 2) ? -y : y);
}

RG_HD float UcrtAtan2f(float yf, float xf) {
    const double y = static_cast<double>(yf), x = static_cast<double>(xf);
    const uint64_t xb = DoubleBits(x), yb = DoubleBits(y);
    const uint64_t ax = xb & 0x7fffffffffffffffull, ay = yb & 0x7fffffffffffffffull;
    constexpr uint64_t kInf = 0x7ff0000000000000ull;
    const bool yneg = (yb >> 63) != 0, xneg = (xb >> 63) != 0;
    const int exp_gap = static_cast<int>((yb >> 52) & 0x7ff) - static_cast<int>((xb >> 52) & 0x7ff);
    if (ax > kInf) return BitsFloat(FloatBits(xf) | 0x400000u);
    if (ay > kInf) return BitsFloat(FloatBits(yf) | 0x400000u);
    if (ay == 0) return xneg ? BitsFloat(yneg ? 0xc0490fdbu : 0x40490fdbu) : yf;
    if (ax == 0 || exp_gap > 26) return BitsFloat(yneg ? 0xbfc90fdbu : 0x3fc90fdbu);
    if (exp_gap < -13 && !xneg) {
        if (exp_gap < -150) return yneg ? -0.0f : 0.0f;
        return static_cast<float>(y / x);
    }
    if (exp_gap < -26) return BitsFloat(yneg ? 0xc0490fdbu : 0x40490fdbu);
    if (ay == kInf) {
        if (!xneg) return BitsFloat(yneg ? 0xbf490fdbu : 0x3f490fdbu);
        return BitsFloat(yneg ? 0xc016cbe4u : 0x4016cbe4u);
    }
    double lo = fabs(y), hi = fabs(x);
    const bool swapped = lo > hi;
    if (swapped) {
        const double t = lo;
        lo = hi;
        hi = t;
    }
    const double ratio = lo / hi;
    double res;
    if (ratio > 0x1p-4) {
        // atan(ratio) = atan(i/256) + atan(r), r = (256 lo - i hi) / (i lo + 256 hi).
        const int i = static_cast<int>(fma(256.0, ratio, 0.5));
        const double id = static_cast<double>(i);
        const double r = fma(lo, 256.0, -(id * hi)) / fma(lo, id, hi * 256.0);
        res = fma(-((r * r) * r), 0x1.5555555550877p-2, r + MathTables::kAtanTab[i - 16]);
    } else if (ratio < 0x1.a36e2eb1c432dp-14) {
        res = ratio;
    } else {
        const double r2 = ratio * ratio;
        double p = fma(-0x1.2492482bd6be1p-3, r2, 0x1.99999999643a3p-3);
        p = fma(-p, r2, 0x1.5555555555538p-2);
        res = fma(-(r2 * ratio), p, ratio);
    }
    if (swapped) res = 0x1.921fb54442d18p+0 - res;
    if (xneg) res = 0x1.921fb54442d18p+1 - res;
    return static_cast<float>(yneg ? -res : res);
}

// exp(t) with the sign bits `sign` or'ed in; `fused` selects the FMA evaluation of the reduced polynomial.
RG_HD float UcrtPowExp(double t, uint32_t sign, bool fused) {
    if (t > 0x1.62e43p6) return BitsFloat(0x7f800000u | sign);
    if (!(t > -0x1.9d1dap6)) return BitsFloat(sign);
    const int n = static_cast<int>(rint(t * 0x1.71547652b82fep+6));
    const double tab = MathTables::kExp2Tab[n & 63];
    double res;
    if (fused) {
        const double r = fma(-static_cast<double>(n), 0x1.62e42fefa39efp-7, t);
        res = fma(r * fma(fma(0x1.5555555555555p-3, r, 0.5), r, 1.0), tab, tab);
    } else {
        const double r = t - static_cast<double>(n) * 0x1.62e42fefa39efp-7;
        res = ((r * r) * (0x1.5555555555555p-3 * r + 0.5) + r) * tab + tab;
    }
    res = BitsDouble(DoubleBits(res) + (static_cast<uint64_t>(static_cast<int64_t>(n >> 6)) << 52));
    return BitsFloat(FloatBits(static_cast<float>(res)) | sign);
}

RG_HD float UcrtPowf(float x, float y) {
    const uint32_t ux = FloatBits(x), uy = FloatBits(y);
    const uint32_t ax = ux & 0x7fffffffu, ay = uy & 0x7fffffffu;
    const auto odd_int = [&]() {
        if ((ay & 0x7f800000u) > 0x4b000000u) return false;
        const float yr = rintf(y);
        return yr == y && (static_cast<int>(yr) & 1) != 0;
    };
    if (ay >= 0x7f800000u) {
        if (ay > 0x7f800000u) {
            return ux == 0x3f800000u ? (ay < 0x7fc00000u ? 1.5f : 1.0f)
                                     : BitsFloat((ax > 0x7f800000u ? ux : uy) | 0x400000u);
        }
        if (ax > 0x7f800000u) return BitsFloat(ux | 0x400000u);
        if (ax == 0x3f800000u) return 1.0f;
        return BitsFloat((static_cast<int32_t>(uy) < 0) == (ax < 0x3f800000u) ? 0x7f800000u : 0u);
    }
    if (ay == 0) return (ax > 0x7f800000u && ax < 0x7fc00000u) ? BitsFloat(ux | 0x400000u) : 1.0f;
    if (uy == 0x3f800000u) return ax > 0x7f800000u ? BitsFloat(ux | 0x400000u) : x;
    if (ax > 0x7f800000u) return BitsFloat(ux | 0x400000u);
    if (ax == 0x7f800000u || ax == 0) {
        const uint32_t mag = ((static_cast<int32_t>(uy) < 0) == (ax == 0)) ? 0x7f800000u : 0u;
        return BitsFloat(mag | (odd_int() ? (ux & 0x80000000u) : 0u));
    }
    uint32_t sign = 0;
    if (static_cast<int32_t>(ux) < 0 && (ay & 0x7f800000u) <= 0x4b000000u) {
        if (rintf(y) != y) return BitsFloat(0xffc00000u);
        if (odd_int()) sign = 0x80000000u;
    }
    const double dx = static_cast<double>(x), dy = static_cast<double>(y);
    if (static_cast<int32_t>(ux) < 0x3f880000) {
        const double f = dx - 1.0;
        if (fabs(f) < 0x1p-4) {
            // log1p(f) by the atanh series in u = f / (f + 2).
            const double u = f / (f + 2.0);
            const double w = u + u;
            const double v = w * w;
            const double w3 = w * v;
            const double a = v * 0x1.9999999bac6d4p-7 + 0x1.55555555554e6p-4;
            const double b = v * 0x1.c8034c85dfff0p-12 + 0x1.2492307f1519fp-9;
            const double s = (((v * v) * w3) * b + w3 * a) - u * f;
            return UcrtPowExp(dy * (f + s), sign, false);
        }
    }
    const uint64_t bx = DoubleBits(dx);
    const uint64_t mant = bx & 0x000fffffffffffffull;
    const uint32_t idx = static_cast<uint32_t>(mant >> 44) + static_cast<uint32_t>((mant >> 43) & 1);
    const double c = BitsDouble(static_cast<uint64_t>(0x3fe00u | idx) << 44);
    const double m = BitsDouble(mant | 0x3fe0000000000000ull);
    const int e = static_cast<int>((bx >> 52) & 0x7ff) - 0x3ff;
    const double r = (c - m) * MathTables::kInvCTab[idx];
    const double rq = r * fma(fma(0x1.5555555555555p-2, r, 0.5), r, 1.0);
    const double log_x = (static_cast<double>(e) * 0x1.62e42fefa39efp-1 + MathTables::kLogCTab[idx]) - rq;
    return UcrtPowExp(dy * log_x, sign, true);
}

RG_HD float UcrtAcosf(float x) {
    const uint32_t ux = FloatBits(x);
    const uint32_t ax = ux & 0x7fffffffu, ex = (ux >> 23) & 0xff;
    if (ax > 0x7f800000u) return BitsFloat(ux | 0x400000u);
    if (ex < 0x65) return BitsFloat(0x3fc90fdbu);
    if (ex >= 0x7f) {
        if (x == 1.0f) return 0.0f;
        if (x == -1.0f) return BitsFloat(0x40490fdbu);
        return BitsFloat(0xffc00000u);
    }
    const float a = fabsf(x);
    const bool big = ex >= 0x7e;
    const float z = big ? (1.0f - a) * 0.5f : a * a;
    const float s = big ? sqrtf(z) : 0.0f;
    float t = BitsFloat(0x3b81ce6bu);
    t = fmaf(-t, z, BitsFloat(0xbc5b3fe1u));
    t = fmaf(t, z, -BitsFloat(0x3d678bddu));
    t = fmaf(t, z, BitsFloat(0x3e3c94dcu));
    const float r = (t * z) / fmaf(-BitsFloat(0x3f561f0du), z, BitsFloat(0x3f8d6fa5u));
    if (!big) {
        const double lo = 0x1.1a62633145c07p-54 - static_cast<double>(r * x);
        return static_cast<float>(0x1.921fb54442d18p+0 - (static_cast<double>(x) - lo));
    }
    if (static_cast<int32_t>(ux) < 0) {
        const double h = (static_cast<double>(s * r) - 0x1.1a62633145c07p-54) + static_cast<double>(s);
        return static_cast<float>(0x1.921fb54442d18p+1 - (h + h));
    }
    const float hi = BitsFloat(FloatBits(s) & 0xffff0000u);
    const float c = fmaf(-hi, hi, z) / (s + hi);
    return fmaf(r, s + s, c + c) + (hi + hi);
}

// ------------------------------------------------------------------------------------------------------------
// sinf / cosf over the whole float range, as the reference computes them (RocketGpuSinf / RocketGpuCosf,
// math.hpp:494-586): the UCRT kernels below |x| = 0x4b800456 (sinf) / 0x4f490fdb (cosf); above, the host C
// library's sinf / cosf, which on the reference's platform is glibc's (sysdeps/ieee754/flt-32/s_sinf.c,
// s_cosf.c, sincosf.h and sincosf_data.c, glibc 2.28 and later, in the FMA build glibc selects on FMA hardware).
// That algorithm, restated: reduce |x| modulo pi/2 exactly with a 96-bit window of 2/pi picked by the exponent
// (reduce_large), then evaluate an even (cos) or odd (sin) double polynomial of the reduced argument, the
// quadrant choosing which and its sign.  Infinities and NaNs give a NaN.

// glibc's __sincosf_table: the quadrant signs and the polynomial coefficients for sin (s1..s3) and cos (c0..c4),
// entry 1 with the cos coefficients negated (quadrants 2 and 3).
struct LibmSincosCoeffs {
    double sign[4];
    double c0, c1, c2, c3, c4;
    double s1, s2, s3;
};
struct LibmSincosTable {
    static constexpr LibmSincosCoeffs k[2] = {
        {{1.0, -1.0, -1.0, 1.0}, 0x1p0, -0x1.ffffffd0c621cp-2, 0x1.55553e1068f19p-5, -0x1.6c087e89a359dp-10,
         0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7, -0x1.994eb3774cf24p-13},
        {{1.0, -1.0, -1.0, 1.0}, -0x1p0, 0x1.ffffffd0c621cp-2, -0x1.55553e1068f19p-5, 0x1.6c087e89a359dp-10,
         -0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7, -0x1.994eb3774cf24p-13}};
};
RG_HD const LibmSincosCoeffs& LibmSincos(int negated_cos) { return LibmSincosTable::k[negated_cos]; }

// x in [-pi/4, pi/4], x2 = x * x: sin x for even n, cos x for odd n (glibc's sinf_poly with FMA).
RG_HD float LibmSincosPoly(double x, double x2, const LibmSincosCoeffs& p, int n) {
    if ((n & 1) == 0) {
        const double x3 = x * x2;
        const double tail = fma(x2, p.s3, p.s2);
        const double x7 = x3 * x2;
        const double head = fma(x3, p.s1, x);
        return static_cast<float>(fma(x7, tail, head));
    }
    const double x4 = x2 * x2;
    const double tail = fma(x2, p.c4, p.c3);
    const double head = fma(x2, p.c1, p.c0);
    const double x6 = x4 * x2;
    const double mid = fma(x4, p.c2, head);
    return static_cast<float>(fma(x6, tail, mid));
}

// |x| modulo pi/2 for |x| >= 2^24 (glibc's reduce_large): the 24-bit significand times the 96 bits of 2/pi that
// matter at this exponent gives the quadrant n in the top bits and the reduced argument, scaled by pi * 2^-63.
RG_HD double LibmReduceLarge(uint32_t bits, int* quadrant) {
    const uint32_t* w = &MathTables::kTwoOverPiBits[(bits >> 26) & 15];
    const int shift = static_cast<int>((bits >> 23) & 7);
    const uint32_t m = ((bits & 0xffffffu) | 0x800000u) << shift;
    uint64_t hi = static_cast<uint32_t>(m * w[0]);
    const uint64_t mid = static_cast<uint64_t>(m) * w[4];
    const uint64_t lo = static_cast<uint64_t>(m) * w[8];
    hi = (lo >> 32) | (hi << 32);
    hi += mid;
    const uint64_t n = (hi + (1ull << 61)) >> 62;
    hi -= n << 62;
    *quadrant = static_cast<int>(n);
    return static_cast<double>(static_cast<int64_t>(hi)) * 0x1.921fb54442d18p-62;
}

RG_HD float GameSinf(float x) {
    const uint32_t bits = FloatBits(x);
    if ((bits & 0x7fffffffu) < kUcrtSinfLimitBits) return UcrtSinf(x);
    if ((bits & 0x7f800000u) == 0x7f800000u) return x - x;  // inf or NaN: NaN
    int n = 0;
    const double r = LibmReduceLarge(bits, &n);
    const int q = n + static_cast<int>(bits >> 31);
    return LibmSincosPoly(r * LibmSincos(0).sign[q & 3], r * r, LibmSincos((q & 2) ? 1 : 0), n);
}

RG_HD float GameCosf(float x) {
    const uint32_t bits = FloatBits(x);
    if ((bits & 0x7fffffffu) < kUcrtCosfLimitBits) return UcrtCosf(x);
    if ((bits & 0x7f800000u) == 0x7f800000u) return x - x;
    int n = 0;
    const double r = LibmReduceLarge(bits, &n);
    return LibmSincosPoly(r * LibmSincos(0).sign[n & 3], r * r, LibmSincos((n & 2) ? 1 : 0), n ^ 1);
}

// atan2 in double rounded to float (SafeAtan2, math.hpp:1546-1552, the car-car culling's angle).  The device's
// double atan2 and the host C library's agree to within an ulp of double, so the float results agree except when a
// double result lies within that ulp of a float rounding boundary.
RG_HD float SafeAtan2(float y, float x) {
    return static_cast<float>(atan2(static_cast<double>(y), static_cast<double>(x)));
}

// Rotation by axis/angle as Bullet's btQuaternion::setRotation (math.hpp:876-880).
RG_HD Quat QuatFromAxisAngle(const Vec3& axis, float angle) {
    const float d = Length(axis);
    const float s = GameSinf(angle * 0.5f) / d;
    return Quat(GameCosf(angle * 0.5f), axis.x * s, axis.y * s, axis.z * s);
}

// q * v * conj(q) in Bullet's btQuatRotate operation order (math.hpp:882-895).
RG_HD Vec3 QuatRotate(const Quat& q, const Vec3& v) {
    const float rx = q.w * v.x + q.y * v.z - q.z * v.y;
    const float ry = q.w * v.y + q.z * v.x - q.x * v.z;
    const float rz = q.w * v.z + q.x * v.y - q.y * v.x;
    const float rw = -q.x * v.x - q.y * v.y - q.z * v.z;
    const float cx = -q.x, cy = -q.y, cz = -q.z, cw = q.w;
    return Vec3(rw * cx + rx * cw + ry * cz - rz * cy, rw * cy + ry * cw + rz * cx - rx * cz,
                rw * cz + rz * cw + rx * cy - ry * cx);
}

}  // namespace core
}  // namespace rg
