// Scalar/vector math, PRNG and sampling warps for the fhp path-tracing core.
// __host__ __device__ so the host-side builders and the gfx950 kernels share one definition.
// All arithmetic is written in the reference's operation order and compiled with
// -ffp-contract=off (no implicit FMA), which keeps FP32 results reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FHP_HD __host__ __device__ __forceinline__

// Per-path device code (traversal, the WARDiffPathTracer paths) is __device__ in the library.
// The host check build (FHP_HOST_CHECK, a CPU test build under AddressSanitizer) compiles the
// same code for the CPU, with single-lane traversal stacks and plain gradient adds.
#ifdef FHP_HOST_CHECK
#define FHP_DEV __host__ __device__
#define FHP_DEVI __host__ __device__ __forceinline__
#else
#define FHP_DEV __device__
#define FHP_DEVI __device__ __forceinline__
#endif

namespace fhp {

struct f3 {
    float x, y, z;
};
FHP_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
FHP_HD f3 mk3(float s) { return f3{s, s, s}; }
FHP_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
FHP_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FHP_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
FHP_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
FHP_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
FHP_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
FHP_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
FHP_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FHP_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
FHP_HD float length(f3 a) { return sqrtf(dot(a, a)); }
FHP_HD f3 normalize(f3 a) { float l = length(a); return a / l; }
FHP_HD bool anyPos(f3 a) { return a.x > 0.f || a.y > 0.f || a.z > 0.f; }
FHP_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
FHP_HD float fminf_(float a, float b) { return a < b ? a : b; }
FHP_HD float fmaxf_(float a, float b) { return a > b ? a : b; }
FHP_HD float saturate(float v) { return fminf_(fmaxf_(v, 0.f), 1.f); }
// Hardware square root, reciprocal square root and reciprocal (about 1 ulp) on the device, for
// heuristics whose exact rounding does not matter (both the sampling and the pdf evaluation use
// the same function); the host build uses the correctly rounded operations.
FHP_HD float hwSqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
FHP_HD float hwRsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.f / sqrtf(x);
#endif
}
FHP_HD float hwRcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.f / x;
#endif
}
FHP_HD float luminance(f3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }  // ColorHelpers.slang:45-49

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kPi2 = 1.57079632679489661923f;
constexpr float kPi4 = 0.785398163397448309616f;
constexpr float kTwoPi = 6.28318530717958647693f;
constexpr float kMinCosTheta = 1e-6f;    // IBSDF.slang:35
constexpr float kMinGGXAlpha = 0.0064f;  // StandardBSDF.slang:53
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kInf = __builtin_huge_valf();
constexpr float kFltMin = 1.175494351e-38f;
constexpr uint32_t kInvalid = 0xffffffffu;

FHP_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
FHP_HD uint32_t popc32(uint32_t x) { return uint32_t(__builtin_popcount(x)); }
// Gradient accumulation into a hash-grid row: a float atomic on the GPU, a plain add in the
// single-threaded host check build.
FHP_DEVI void gradAdd(float* p, float v) {
#ifdef FHP_HOST_CHECK
    *p += v;
#else
    atomicAdd(p, v);
#endif
}
FHP_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// ------------------------------------------------------------------ PRNG
// Morton interleave of the low 16 bits (BitTricks.slang:47-62).
FHP_HD uint32_t morton2(uint32_t vx, uint32_t vy) {
    uint32_t x = vx & 0x0000ffffu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    uint32_t y = vy & 0x0000ffffu;
    y = (y | (y << 8)) & 0x00FF00FFu;
    y = (y | (y << 4)) & 0x0F0F0F0Fu;
    y = (y | (y << 2)) & 0x33333333u;
    y = (y | (y << 1)) & 0x55555555u;
    return x | (y << 1);
}
// SplitMix64 step (SplitMix64.slang:56-62).
FHP_HD uint64_t splitmix(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// xoshiro128** state (16 bytes per path); seeded like UniformSampleGenerator.slang:53-72.
struct Rng {
    uint32_t s0, s1, s2, s3;
    FHP_HD uint32_t next() {
        // Xoshiro.slang:54-69
        uint32_t x = s0 * 5u;
        const uint32_t r = ((x << 7) | (x >> 25)) * 9u;
        const uint32_t t = s1 << 9;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = (s3 << 11) | (s3 >> 21);
        return r;
    }
    // SampleGeneratorInterface.slang:50-55
    FHP_HD float u1() { return float(next() >> 8) * 0x1p-24f; }
};
FHP_HD Rng rngSeed(uint32_t px, uint32_t py, uint32_t sampleNumber) {
    uint64_t st = (uint64_t(sampleNumber) << 32) | uint64_t(morton2(px, py));
    uint64_t a = splitmix(st), b = splitmix(st);
    return Rng{uint32_t(a), uint32_t(a >> 32), uint32_t(b), uint32_t(b >> 32)};
}
// Jenkins hash (HashUtils.slang:39-48), used for gradient hash-grid rows.
FHP_HD uint32_t jenkins(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

// ------------------------------------------------------------------ geometry helpers
// Self-intersection offset in integer ULPs along the normal (GeometryHelpers.slang:54-68).
FHP_HD float offsetComp(float p, float n) {
    const int32_t iOff = int32_t(n * (3 * 256.f));
    const float iPos = u2f(uint32_t(int32_t(f2u(p)) + (p < 0.f ? -iOff : iOff)));
    const float fPos = p + n * (3.f / 65536.f);
    return fabsf(p) < (1.f / 16.f) ? fPos : iPos;
}
FHP_HD f3 rayOriginOffset(f3 p, f3 n) { return mk3(offsetComp(p.x, n.x), offsetComp(p.y, n.y), offsetComp(p.z, n.z)); }

// MathHelpers.slang:540-552
FHP_HD f3 perpStark(f3 u) {
    float ax = fabsf(u.x), ay = fabsf(u.y), az = fabsf(u.z);
    uint32_t uyx = (ax - ay) < 0.f ? 1u : 0u;
    uint32_t uzx = (ax - az) < 0.f ? 1u : 0u;
    uint32_t uzy = (ay - az) < 0.f ? 1u : 0u;
    uint32_t xm = uyx & uzx;
    uint32_t ym = (1u ^ xm) & uzy;
    uint32_t zm = 1u ^ (xm | ym);
    return normalize(cross(u, mk3(float(xm), float(ym), float(zm))));
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013; IntersectionHelpers.slang:118-187).
// On success: t and the weights b1, b2 of v1, v2. The caller applies the t range.
FHP_HD bool watertight(f3 o, f3 d, f3 v0, f3 v1, f3 v2, float& t, float& b1, float& b2) {
    float adx = fabsf(d.x), ady = fabsf(d.y), adz = fabsf(d.z);
    int kz = 0;
    if (ady > adx && ady > adz) kz = 1;
    if (adz > adx && adz > ady) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.f) { int s = kx; kx = ky; ky = s; }
    float dx = comp(d, kx), dy = comp(d, ky);
    float Sx = dx / dz, Sy = dy / dz, Sz = 1.f / dz;
    f3 A = v0 - o, B = v1 - o, C = v2 - o;
    float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
    float Ax = comp(A, kx) - Sx * Akz, Ay = comp(A, ky) - Sy * Akz;
    float Bx = comp(B, kx) - Sx * Bkz, By = comp(B, ky) - Sy * Bkz;
    float Cx = comp(C, kx) - Sx * Ckz, Cy = comp(C, ky) - Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = float((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = float((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = float((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float Az = Sz * Akz, Bz = Sz * Bkz, Cz = Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.f / det;
    t = T * rcp;
    b1 = V * rcp;
    b2 = W * rcp;
    return true;
}

// ------------------------------------------------------------------ sampling warps
// Shirley concentric disk + cosine lift (MathHelpers.slang:321-352).
FHP_HD f3 cosineHemisphere(float ux, float uy, float& pdf) {
    ux = 2.f * ux - 1.f;
    uy = 2.f * uy - 1.f;
    float dx, dy;
    if (ux == 0.f && uy == 0.f) { dx = ux; dy = uy; }
    else {
        float phi, r;
        if (fabsf(ux) > fabsf(uy)) { r = ux; phi = (uy / ux) * kPi4; }
        else { r = uy; phi = kPi2 - (ux / uy) * kPi4; }
        dx = r * cosf(phi);
        dy = r * sinf(phi);
    }
    float z = sqrtf(fmaxf_(0.f, 1.f - (dx * dx + dy * dy)));
    pdf = z * kInvPi;
    return mk3(dx, dy, z);
}
// Uniform triangle barycentrics (MathHelpers.slang:389-394).
FHP_HD f3 triangleSample(float ux, float uy) {
    float su = sqrtf(ux);
    float bx = 1.f - su, by = uy * su;
    return mk3(1.f - bx - by, bx, by);
}

// ------------------------------------------------------------------ microfacet terms
FHP_HD float pow5(float x) { return powf(x, 5.f); }
// Fresnel.slang:42-51
FHP_HD float schlick1(float f0, float f90, float c) { return f0 + (f90 - f0) * pow5(fmaxf_(1.f - c, 0.f)); }
FHP_HD f3 schlick3(f3 f0, float f90, float c) {
    float w = pow5(fmaxf_(1.f - c, 0.f));
    return mk3(f0.x + (f90 - f0.x) * w, f0.y + (f90 - f0.y) * w, f0.z + (f90 - f0.z) * w);
}
// IsotropicGGX.slang:46-51
FHP_HD float ggxD(float alpha, float c) {
    float a2 = alpha * alpha;
    float d = ((c * a2 - c) * c + 1.f);
    return a2 / (d * d * kPi);
}
// IsotropicGGX.slang:148-154
FHP_HD float ggxG1(float a2, float c) {
    if (c <= 0.f) return 0.f;
    float c2 = c * c;
    float t2 = fmaxf_(1.f - c2, 0.f) / c2;
    return 2.f / (1.f + sqrtf(1.f + a2 * t2));
}
// IsotropicGGX.slang:163-169
FHP_HD float ggxLambda(float a2, float c) {
    if (c <= 0.f) return 0.f;
    float c2 = c * c;
    float t2 = fmaxf_(1.f - c2, 0.f) / c2;
    return 0.5f * (-1.f + sqrtf(1.f + a2 * t2));
}
// IsotropicGGX.slang:200-206 (height-correlated Smith)
FHP_HD float ggxG(float alpha, float ci, float co) {
    float a2 = alpha * alpha;
    float li = ggxLambda(a2, ci), lo = ggxLambda(a2, co);
    return 1.f / (1.f + li + lo);
}
// IsotropicGGX.slang:96-101
FHP_HD float ggxPdfVndf(float alpha, f3 wi, f3 h) {
    float G1 = ggxG1(alpha * alpha, wi.z);
    float D = ggxD(alpha, h.z);
    return G1 * D * fmaxf_(0.f, dot(wi, h)) / wi.z;
}
// Spherical-cap VNDF sampling, Dupuy & Benyoub 2023 (IsotropicGGX.slang:103-138).
FHP_HD f3 ggxSampleVndf(float alpha, f3 wi, float ux, float uy, float& pdf) {
    f3 Vh = normalize(mk3(alpha * wi.x, alpha * wi.y, wi.z));
    float phi = 2.f * kPi * ux;
    float z = fmaf((1.f - uy), (1.f + Vh.z), -Vh.z);
    float sinT = sqrtf(saturate(1.f - z * z));
    f3 Nh = mk3(sinT * cosf(phi), sinT * sinf(phi), z) + Vh;
    f3 h = normalize(mk3(alpha * Nh.x, alpha * Nh.y, Nh.z));
    pdf = ggxPdfVndf(alpha, wi, h);
    return h;
}

}  // namespace fhp
