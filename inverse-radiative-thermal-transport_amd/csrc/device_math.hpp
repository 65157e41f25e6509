// device_math.hpp -- FP32 photon-tracing arithmetic for gfx950.
//
// Every function follows the reference shader's operation order (cited per
// function, paths relative to the reference repository root).  The library is
// compiled with -ffp-contract=off and correctly rounded FP32 division and
// square root, so each expression rounds exactly like the scalar CPU
// restatement used to check it; no fused multiply-add is formed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace irtt {

struct V3 {
    float x, y, z;
};

// Correctly rounded FP32 square root and division: the library is built with
// -fhip-fp32-correctly-rounded-divide-sqrt, under which llvm.sqrt.f32 and fdiv
// lower to IEEE round-to-nearest sequences.
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float div_rn(float a, float b) { return a / b; }

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// GLSL normalize(v) = v / length(v); the length is returned through len
__device__ __forceinline__ V3 normalize3(V3 v, float &len)
{
    len = sqrt_rn(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3{div_rn(v.x, len), div_rn(v.y, len), div_rn(v.z, len)};
}
__device__ __forceinline__ V3 normalize3(V3 v)
{
    float l;
    return normalize3(v, l);
}
// p0*b0 + p1*b1 + p2*b2 (thermal_tracing_unified.glsl:139,170)
__device__ __forceinline__ V3 bary_point(V3 p0, V3 p1, V3 p2, float b0, float b1, float b2)
{
    return V3{p0.x * b0 + p1.x * b1 + p2.x * b2, p0.y * b0 + p1.y * b1 + p2.y * b2,
              p0.z * b0 + p1.z * b1 + p2.z * b2};
}
// findMaxIdx (thermal_tracing_unified.glsl:199-211): ties go to the later corner
__device__ __forceinline__ int max_corner(float b0, float b1, float b2)
{
    int idx = 0;
    float v = b0;
    if (b1 >= v) {
        idx = 1;
        v = b1;
    }
    if (b2 >= v) idx = 2;
    return idx;
}

// ---- RNG: assets/shader/utils/glsl/rendering_utils.glsl:361-392 ----
__device__ __forceinline__ uint32_t tea16(uint32_t v0, uint32_t v1)
{
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
__device__ __forceinline__ float rnd(uint32_t &s)
{
    s = 1664525u * s + 1013904223u;
    return __uint_as_float(0x3f800000u | (s >> 9)) - 1.0f;
}

// sin/cos of 2*pi*u: nearest quadrant of 4u, Taylor polynomials of
// sin(pi f/2), cos(pi f/2) on |f| <= 1/2, quadrant rotation.
__device__ __forceinline__ void sincos_turns(float u, float &s, float &c)
{
    float v = u * 4.0f;
    int q = (int)(v + 0.5f);
    float f = v - (float)q;
    float f2 = f * f;
    float ps = 1.604411809e-04f;
    ps = ps * f2 + -4.681753926e-03f;
    ps = ps * f2 + 7.969262451e-02f;
    ps = ps * f2 + -6.459640861e-01f;
    ps = ps * f2 + 1.570796371e+00f;
    float sf = ps * f;
    float pc = -2.520204180e-05f;
    pc = pc * f2 + 9.192602593e-04f;
    pc = pc * f2 + -2.086348087e-02f;
    pc = pc * f2 + 2.536695004e-01f;
    pc = pc * f2 + -1.233700514e+00f;
    float cf = pc * f2 + 1.0f;
    switch (q & 3) {
    case 0: s = sf; c = cf; break;
    case 1: s = cf; c = -sf; break;
    case 2: s = -sf; c = -cf; break;
    default: s = -cf; c = sf; break;
    }
}

// Malley cosine-weighted hemisphere via the uniform disk
// (ray_tracing_utils.glsl:178-182,217-220)
__device__ __forceinline__ V3 cosine_hemisphere(float u0, float u1)
{
    float r = sqrt_rn(u0);
    float s, c;
    sincos_turns(u1, s, c);
    float x = r * c, y = r * s;
    float zz = 1.0f - x * x - y * y;
    return V3{x, y, sqrt_rn(zz > 0.0f ? zz : 0.0f)};
}

// getPerpendicularVector + tangentSpaceToWorldSpace
// (rendering_utils.glsl:212-219,261-272)
__device__ __forceinline__ V3 tangent_to_world(V3 v, V3 n)
{
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    uint32_t xm = ((ax - ay) < 0.0f && (ax - az) < 0.0f) ? 1u : 0u;
    uint32_t ym = (ay - az) < 0.0f ? (1u ^ xm) : 0u;
    uint32_t zm = 1u ^ (xm | ym);
    V3 t = normalize3(cross3(n, V3{(float)xm, (float)ym, (float)zm}));
    V3 b = normalize3(cross3(n, t));
    return V3{t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z,
              t.z * v.x + b.z * v.y + n.z * v.z};
}

// Ray Tracing Gems ch. 6 self-intersection offset (ray_tracing_utils.glsl:242-259)
__device__ __forceinline__ float offset1(float p, float n)
{
    int32_t of = (int32_t)(256.0f * n);
    int32_t bits = __float_as_int(p) + ((p < 0.0f) ? -of : of);
    return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : __int_as_float(bits);
}
__device__ __forceinline__ V3 offset_ray(V3 p, V3 n)
{
    return V3{offset1(p.x, n.x), offset1(p.y, n.y), offset1(p.z, n.z)};
}

// ---- watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), two-sided ----
__device__ __forceinline__ float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ RayPre ray_pre(V3 d)
{
    RayPre r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    float am = ax;
    if (ay > am) {
        kz = 1;
        am = ay;
    }
    if (az > am) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    float dz = comp(d, kz);
    if (dz < 0.0f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = div_rn(comp(d, kx), dz);
    r.Sy = div_rn(comp(d, ky), dz);
    r.Sz = div_rn(1.0f, dz);
    return r;
}

__device__ __forceinline__ bool tri_hit(const RayPre &rp, V3 o, V3 p0, V3 p1, V3 p2, float tmin,
                                        float tmax, float &t_out, float &b1, float &b2)
{
    V3 A = p0 - o, B = p1 - o, C = p2 - o;
    float Akz = comp(A, rp.kz), Bkz = comp(B, rp.kz), Ckz = comp(C, rp.kz);
    float Ax = comp(A, rp.kx) - rp.Sx * Akz, Ay = comp(A, rp.ky) - rp.Sy * Akz;
    float Bx = comp(B, rp.kx) - rp.Sx * Bkz, By = comp(B, rp.ky) - rp.Sy * Bkz;
    float Cx = comp(C, rp.kx) - rp.Sx * Ckz, Cy = comp(C, rp.ky) - rp.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (rp.Sz * Akz) + V * (rp.Sz * Bkz) + W * (rp.Sz * Ckz);
    float t = div_rn(T, det);
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    b1 = div_rn(V, det);
    b2 = div_rn(W, det);
    return true;
}

}  // namespace irtt
