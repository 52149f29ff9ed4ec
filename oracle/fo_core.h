// Oracle core: a scalar CPU restatement of the reference path tracer's algorithms.
//
// TEST INFRASTRUCTURE ONLY. This code is the checker that the HIP product is compared
// against; it is never linked into or called by the product library. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Every function cites the reference file:line it restates. Arithmetic follows the
// reference's operation order with FP contraction disabled (-ffp-contract=off), so
// that FP32 results are reproducible against the GPU path built the same way.
// Templates over T let the integrators also run in FP64 for converged-RMSE checks.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <limits>
#include <vector>

namespace fo {

// ---------------------------------------------------------------- vector math
template <typename T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T a) : x(a), y(a), z(a) {}
    template <typename U> explicit V3(const V3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// normalize(v) = v / length(v), IEEE division (the product uses the same definition).
template <typename T> inline V3<T> normalize(V3<T> a) { T l = length(a); return a / l; }
template <typename T> inline bool anyPositive(V3<T> a) { return a.x > T(0) || a.y > T(0) || a.z > T(0); }
template <typename T> inline T luminance(V3<T> c) {
    // ColorHelpers.slang:45-49
    return c.x * T(0.2126f) + c.y * T(0.7152f) + c.z * T(0.0722f);
}
template <typename T> inline T saturate(T v) { return std::min(std::max(v, T(0)), T(1)); }

template <typename T> struct Consts;
template <> struct Consts<float> {
    static constexpr float pi = 3.14159265358979323846f, inv_pi = 0.318309886183790671538f;
    static constexpr float pi_2 = 1.57079632679489661923f, pi_4 = 0.785398163397448309616f;
    static constexpr float two_pi = 6.28318530717958647693f;
};
template <> struct Consts<double> {
    static constexpr double pi = 3.14159265358979323846, inv_pi = 0.318309886183790671538;
    static constexpr double pi_2 = 1.57079632679489661923, pi_4 = 0.785398163397448309616;
    static constexpr double two_pi = 6.28318530717958647693;
};

inline uint32_t asuint(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float asfloat(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ---------------------------------------------------------------- PRNG
// BitTricks.slang:47-62 (interleave_32bit, Morton code of the low 16 bits of x and y).
inline uint32_t interleave_32bit(uint32_t vx, uint32_t vy) {
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
// SplitMix64.slang:56-62 (Steele et al. SplitMix64, public constants).
inline uint64_t splitmix64_next(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
// Xoshiro.slang:54-69 (Blackman-Vigna xoshiro128**).
struct SampleGen {
    uint32_t s[4];
    uint32_t next() {
        const uint32_t result = rotl32(s[0] * 5u, 7) * 9u;
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl32(s[3], 11);
        return result;
    }
    // SampleGeneratorInterface.slang:50-55: upper 24 bits / 2^24.
    float next1D() { return float(next() >> 8) * 0x1p-24f; }
};
// UniformSampleGenerator.slang:64-69: SplitMix64(state = interleave(pixel) | sampleNumber<<32),
// two 64-bit draws form the xoshiro128** state.
inline SampleGen makeSampleGen(uint32_t px, uint32_t py, uint32_t sampleNumber) {
    uint64_t state = (uint64_t(sampleNumber) << 32) | uint64_t(interleave_32bit(px, py));
    uint64_t s0 = splitmix64_next(state);
    uint64_t s1 = splitmix64_next(state);
    SampleGen g;
    g.s[0] = uint32_t(s0); g.s[1] = uint32_t(s0 >> 32); g.s[2] = uint32_t(s1); g.s[3] = uint32_t(s1 >> 32);
    return g;
}
// HashUtils.slang:39-48 (Jenkins 32-bit integer hash).
inline uint32_t jenkinsHash(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

// ---------------------------------------------------------------- geometry helpers
// GeometryHelpers.slang:54-68 (Waechter-Binder self-intersection offset), FP32 only.
inline V3<float> computeRayOrigin(V3<float> pos, V3<float> n) {
    const float origin = 1.f / 16.f;
    const float fScale = 3.f / 65536.f;
    const float iScale = 3 * 256.f;
    V3<float> r;
    for (int i = 0; i < 3; i++) {
        int32_t iOff = int32_t(n[i] * iScale);
        float p = pos[i];
        float iPos = asfloat(uint32_t(int32_t(asuint(p)) + (p < 0.f ? -iOff : iOff)));
        float fOff = n[i] * fScale;
        r[i] = std::fabs(p) < origin ? p + fOff : iPos;
    }
    return r;
}
inline V3<double> computeRayOrigin(V3<double> pos, V3<double> n) {
    // FP64 oracle: offset the FP32-rounded point, then widen (keeps the same escape distance).
    V3<float> r = computeRayOrigin(V3<float>(pos), V3<float>(n));
    return V3<double>(r);
}

// MathHelpers.slang:540-552
template <typename T> inline V3<T> perp_stark(V3<T> u) {
    V3<T> a(std::fabs(u.x), std::fabs(u.y), std::fabs(u.z));
    uint32_t uyx = (a.x - a.y) < T(0) ? 1 : 0;
    uint32_t uzx = (a.x - a.z) < T(0) ? 1 : 0;
    uint32_t uzy = (a.y - a.z) < T(0) ? 1 : 0;
    uint32_t xm = uyx & uzx;
    uint32_t ym = (1 ^ xm) & uzy;
    uint32_t zm = 1 ^ (xm | ym);
    return normalize(cross(u, V3<T>(T(xm), T(ym), T(zm))));
}

// IntersectionHelpers.slang:118-187 (Woop-Benthin-Wald watertight test).
// Returns true with t and weights (b0,b1,b2) of v0,v1,v2; the caller checks the t range.
template <typename T>
inline bool intersectRayTriangle(V3<T> o, V3<T> d, const V3<T> v[3], V3<T>& bary, T& t) {
    V3<T> ad(std::fabs(d.x), std::fabs(d.y), std::fabs(d.z));
    int axis = 0;
    if (ad.y > ad.x && ad.y > ad.z) axis = 1;
    if (ad.z > ad.x && ad.z > ad.y) axis = 2;
    int kz = axis, kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (d[kz] < T(0)) std::swap(kx, ky);
    T Sx = d[kx] / d[kz], Sy = d[ky] / d[kz], Sz = T(1) / d[kz];
    V3<T> A = v[0] - o, B = v[1] - o, C = v[2] - o;
    T Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    T Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    T Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    T U = Cx * By - Cy * Bx;
    T V = Ax * Cy - Ay * Cx;
    T W = Bx * Ay - By * Ax;
    if (U == T(0) || V == T(0) || W == T(0)) {
        U = T((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = T((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = T((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < T(0) || V < T(0) || W < T(0)) && (U > T(0) || V > T(0) || W > T(0))) return false;
    T det = U + V + W;
    if (det == T(0)) return false;
    T Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    T Tt = U * Az + V * Bz + W * Cz;
    T rcpDet = T(1) / det;
    t = Tt * rcpDet;
    bary = V3<T>(U * rcpDet, V * rcpDet, W * rcpDet);
    return true;
}

// GeometryHelpers.slang:442-457 (Moeller-Trumbore distance, used by the differentiable hit).
template <typename T>
inline T computeRayDistanceToTriangle(const V3<T> v[3], V3<T> o, V3<T> d) {
    V3<T> e1 = v[1] - v[0], e2 = v[2] - v[0];
    V3<T> pvec = cross(d, e2);
    T det = dot(e1, pvec);
    V3<T> tvec = o - v[0];
    V3<T> qvec = cross(tvec, e1);
    return dot(e2, qvec) / det;
}
// GeometryHelpers.slang:465-482
template <typename T>
inline V3<T> computeBarycentrics(const V3<T> v[3], V3<T> p) {
    V3<T> e0 = v[0] - p, e1 = v[1] - p, e2 = v[2] - p;
    T a0 = length(cross(e1, e2)), a1 = length(cross(e2, e0)), a2 = length(cross(e0, e1));
    T a = a0 + a1 + a2;
    return V3<T>(a0 / a, a1 / a, a2 / a);
}

// ---------------------------------------------------------------- sampling warps
// MathHelpers.slang:321-341
template <typename T> inline void sample_disk_concentric(T ux, T uy, T& dx, T& dy) {
    ux = T(2) * ux - T(1);
    uy = T(2) * uy - T(1);
    if (ux == T(0) && uy == T(0)) { dx = ux; dy = uy; return; }
    T phi, r;
    if (std::fabs(ux) > std::fabs(uy)) { r = ux; phi = (uy / ux) * Consts<T>::pi_4; }
    else { r = uy; phi = Consts<T>::pi_2 - (ux / uy) * Consts<T>::pi_4; }
    dx = r * std::cos(phi);
    dy = r * std::sin(phi);
}
// MathHelpers.slang:346-352
template <typename T> inline V3<T> sample_cosine_hemisphere_concentric(T ux, T uy, T& pdf) {
    T dx, dy;
    sample_disk_concentric(ux, uy, dx, dy);
    T z = std::sqrt(std::max(T(0), T(1) - (dx * dx + dy * dy)));
    pdf = z * Consts<T>::inv_pi;
    return V3<T>(dx, dy, z);
}
// MathHelpers.slang:389-394
template <typename T> inline V3<T> sample_triangle(T ux, T uy) {
    T su = std::sqrt(ux);
    T bx = T(1) - su, by = uy * su;
    return V3<T>(T(1) - bx - by, bx, by);
}
// MathHelpers.slang:403-411 (Jakob 2012 vMF inversion).
template <typename T> inline V3<T> sampleVonMisesFisher(T ux, T uy, T kappa) {
    T sy = std::max(T(1) - uy, T(1e-6f));
    T phi = Consts<T>::two_pi * ux;
    T cosTheta = T(1) + std::log(std::exp(T(-2) * kappa) * (T(1) - sy) + sy) / kappa;
    T sinTheta = std::sqrt(T(1) - cosTheta * cosTheta);
    return V3<T>(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);
}

// ---------------------------------------------------------------- BSDFs
constexpr float kMinCosTheta = 1e-6f;   // IBSDF.slang:35
constexpr float kMinGGXAlpha = 0.0064f; // StandardBSDF.slang:53

// Fresnel.slang:42-51
template <typename T> inline T fresnelSchlick(T f0, T f90, T c) {
    return f0 + (f90 - f0) * std::pow(std::max(T(1) - c, T(0)), T(5));
}
template <typename T> inline V3<T> fresnelSchlick3(V3<T> f0, T f90, T c) {
    T w = std::pow(std::max(T(1) - c, T(0)), T(5));
    return V3<T>(f0.x + (f90 - f0.x) * w, f0.y + (f90 - f0.y) * w, f0.z + (f90 - f0.z) * w);
}
// IsotropicGGX.slang:46-51
template <typename T> inline T evalNdfGGX(T alpha, T cosTheta) {
    T a2 = alpha * alpha;
    T d = ((cosTheta * a2 - cosTheta) * cosTheta + T(1));
    return a2 / (d * d * Consts<T>::pi);
}
// IsotropicGGX.slang:148-154
template <typename T> inline T evalG1GGX(T alphaSqr, T cosTheta) {
    if (cosTheta <= T(0)) return T(0);
    T c2 = cosTheta * cosTheta;
    T tan2 = std::max(T(1) - c2, T(0)) / c2;
    return T(2) / (T(1) + std::sqrt(T(1) + alphaSqr * tan2));
}
// IsotropicGGX.slang:163-169
template <typename T> inline T evalLambdaGGX(T alphaSqr, T cosTheta) {
    if (cosTheta <= T(0)) return T(0);
    T c2 = cosTheta * cosTheta;
    T tan2 = std::max(T(1) - c2, T(0)) / c2;
    return T(0.5f) * (T(-1) + std::sqrt(T(1) + alphaSqr * tan2));
}
// IsotropicGGX.slang:200-206
template <typename T> inline T evalMaskingSmithGGXCorrelated(T alpha, T ci, T co) {
    T a2 = alpha * alpha;
    T li = evalLambdaGGX(a2, ci), lo = evalLambdaGGX(a2, co);
    return T(1) / (T(1) + li + lo);
}
// IsotropicGGX.slang:96-101
template <typename T> inline T evalPdfGGX_VNDF(T alpha, V3<T> wi, V3<T> h) {
    T G1 = evalG1GGX(alpha * alpha, wi.z);
    T D = evalNdfGGX(alpha, h.z);
    return G1 * D * std::max(T(0), dot(wi, h)) / wi.z;
}
// IsotropicGGX.slang:103-138 (Dupuy-Benyoub 2023 spherical-cap VNDF sampling).
template <typename T> inline V3<T> sampleGGX_VNDF(T alpha, V3<T> wi, T ux, T uy, T& pdf) {
    V3<T> Vh = normalize(V3<T>(alpha * wi.x, alpha * wi.y, wi.z));
    T phi = T(2) * Consts<T>::pi * ux;
    // fma((1-u.y),(1+Vh.z),-Vh.z) is an explicit fused multiply-add in the reference.
    T z = std::fma((T(1) - uy), (T(1) + Vh.z), -Vh.z);
    T sinTheta = std::sqrt(saturate(T(1) - z * z));
    T x = sinTheta * std::cos(phi), y = sinTheta * std::sin(phi);
    V3<T> Nh = V3<T>(x, y, z) + Vh;
    V3<T> h = normalize(V3<T>(alpha * Nh.x, alpha * Nh.y, Nh.z));
    pdf = evalPdfGGX_VNDF(alpha, wi, h);
    return h;
}

// Material record shared with the product's scene format (20 floats per material).
enum MatField {
    M_TYPE = 0, M_BASE = 1, M_ROUGH = 4, M_METAL = 5, M_IOR = 6, M_EMIT = 7, M_EMIT_FACTOR = 10,
    M_DOUBLE_SIDED = 11, M_TEX = 12, M_EMIT_TEX = 13,  // M_EMIT_TEX: emissive texture + 1, 0 none
    M_SPEC_TRANS = 14, M_DIFF_TRANS = 15, M_TRANS = 16, M_THIN = 19, M_STRIDE = 20  // M_TRANS: rgb
};
enum MatType { MAT_STANDARD = 0, MAT_PBRT_DIFFUSE = 1 };

// Parameters of one shading point's BSDF (StandardBSDFData: diffuse/specular albedo, roughness,
// metallic, relative IoR eta, transmission albedo and the two transmission weights).
template <typename T> struct BsdfData {
    int type;            // MatType
    V3<T> diffuse, specular;
    T roughness, metallic;
    T eta = T(1);
    V3<T> transmission = V3<T>(T(0));  // after the thin-surface square root
    T specTrans = T(0), diffTrans = T(0);
};

// Fresnel.slang:68-92 (dielectric reflectance; cosT = 0 on total internal reflection).
template <typename T> inline T fresnelDielectric(T eta, T cosI, T& cosT) {
    if (cosI < T(0)) { eta = T(1) / eta; cosI = -cosI; }
    T sin2 = eta * eta * (T(1) - cosI * cosI);
    if (sin2 > T(1)) { cosT = T(0); return T(1); }
    cosT = std::sqrt(T(1) - sin2);
    T rs = (eta * cosI - cosT) / (eta * cosI + cosT);
    T rp = (eta * cosT - cosI) / (eta * cosT + cosI);
    return T(0.5f) * (rs * rs + rp * rp);
}
template <typename T> inline T signOf(T v) { return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0)); }

// StandardBSDF.slang:125-199 (lobe setup) with all four lobes: Frostbite diffuse reflection,
// Lambert diffuse transmission, GGX specular reflection, GGX dielectric reflection/transmission.
// Every lobe is active (Material.cpp:69 sets LobeType::All).
template <typename T> struct StdBsdf {
    V3<T> diffAlbedo, specAlbedo, transAlbedo;
    T roughness, alpha, alphaT, eta, specTrans, diffTrans;
    T pDiff, pDiffT, pSpec, pSpecT;
    void init(V3<T> wi, const BsdfData<T>& d) {
        diffAlbedo = d.diffuse;
        roughness = d.roughness;
        specAlbedo = d.specular;
        transAlbedo = d.transmission;
        eta = d.eta;
        alpha = d.roughness * d.roughness;
        if (alpha < T(kMinGGXAlpha)) alpha = T(0);
        alphaT = d.eta == T(1) ? T(0) : alpha;
        diffTrans = d.diffTrans;
        specTrans = d.specTrans;
        T metallicBRDF = d.metallic * (T(1) - specTrans);
        T dielectricBSDF = (T(1) - d.metallic) * (T(1) - specTrans);
        T specularBSDF = specTrans;
        T diffuseWeight = luminance(d.diffuse);
        T specularWeight = luminance(fresnelSchlick3(d.specular, T(1), wi.z));
        pDiff = diffuseWeight * dielectricBSDF * (T(1) - diffTrans);
        pDiffT = diffuseWeight * dielectricBSDF * diffTrans;
        pSpec = specularWeight * (metallicBRDF + dielectricBSDF);
        pSpecT = specularBSDF;
        T norm = pDiff + pDiffT + pSpec + pSpecT;
        if (norm > T(0)) {
            norm = T(1) / norm;
            pDiff *= norm;
            pDiffT *= norm;
            pSpec *= norm;
            pSpecT *= norm;
        }
    }
    // FrostbiteDiffuseBRDF.slang:84-96 (f * pi)
    V3<T> frostbiteWeight(V3<T> wi, V3<T> wo) const {
        V3<T> h = normalize(wi + wo);
        T woDotH = dot(wo, h);
        T energyBias = T(0) + (T(0.5f) - T(0)) * roughness;               // lerp(0, 0.5, r)
        T energyFactor = T(1) + (T(1) / T(1.51f) - T(1)) * roughness;    // lerp(1, 1/1.51, r)
        T fd90 = energyBias + T(2) * woDotH * woDotH * roughness;
        T wiScatter = fresnelSchlick(T(1), fd90, wi.z);
        T woScatter = fresnelSchlick(T(1), fd90, wo.z);
        return diffAlbedo * wiScatter * woScatter * energyFactor;
    }
    V3<T> evalDiffuse(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, wo.z) < T(kMinCosTheta)) return V3<T>(T(0));
        return frostbiteWeight(wi, wo) * Consts<T>::inv_pi * wo.z;
    }
    // SpecularMicrofacet.slang:58-83
    V3<T> evalSpecular(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, wo.z) < T(kMinCosTheta)) return V3<T>(T(0));
        if (alpha == T(0)) return V3<T>(T(0));
        V3<T> h = normalize(wi + wo);
        T wiDotH = dot(wi, h);
        T D = evalNdfGGX(alpha, h.z);
        T G = evalMaskingSmithGGXCorrelated(alpha, wi.z, wo.z);
        V3<T> F = fresnelSchlick3(specAlbedo, T(1), wiDotH);
        return F * D * G * T(0.25f) / wi.z;
    }
    T pdfDiffuse(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, wo.z) < T(kMinCosTheta)) return T(0);
        return Consts<T>::inv_pi * wo.z;
    }
    // SpecularMicrofacet.slang:163-182
    T pdfSpecular(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, wo.z) < T(kMinCosTheta)) return T(0);
        if (alpha == T(0)) return T(0);
        V3<T> h = normalize(wi + wo);
        T wiDotH = dot(wi, h);
        return evalPdfGGX_VNDF(alpha, wi, h) / (T(4) * wiDotH);
    }
    // LambertDiffuseBTDF.slang:40-84
    V3<T> evalDiffuseT(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, -wo.z) < T(kMinCosTheta)) return V3<T>(T(0));
        return Consts<T>::inv_pi * transAlbedo * -wo.z;
    }
    T pdfDiffuseT(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, -wo.z) < T(kMinCosTheta)) return T(0);
        return Consts<T>::inv_pi * -wo.z;
    }
    // Half vector of SpecularMicrofacetBSDF (SpecularMicrofacet.slang:230-233, :421-423), flipped
    // into the upper hemisphere.
    V3<T> bsdfHalf(V3<T> wi, V3<T> wo, bool isR) const {
        V3<T> h = normalize(wo + wi * (isR ? T(1) : eta));
        return h * signOf(h.z);
    }
    // SpecularMicrofacet.slang:215-256
    V3<T> evalSpecularT(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, std::fabs(wo.z)) < T(kMinCosTheta)) return V3<T>(T(0));
        if (alphaT == T(0)) return V3<T>(T(0));
        bool isR = wo.z > T(0);
        V3<T> h = bsdfHalf(wi, wo, isR);
        T wiDotH = dot(wi, h), woDotH = dot(wo, h);
        T D = evalNdfGGX(alphaT, h.z);
        T G = evalMaskingSmithGGXCorrelated(alphaT, wi.z, std::fabs(wo.z));
        T cosT;
        T F = fresnelDielectric(eta, wiDotH, cosT);
        if (isR) return V3<T>(F * D * G * T(0.25f) / wi.z);
        T sqrtDenom = woDotH + eta * wiDotH;
        T t = eta * eta * wiDotH * woDotH / (wi.z * sqrtDenom * sqrtDenom);
        return transAlbedo * (T(1) - F) * D * G * std::fabs(t);
    }
    // SpecularMicrofacet.slang:412-458
    T pdfSpecularT(V3<T> wi, V3<T> wo) const {
        if (std::min(wi.z, std::fabs(wo.z)) < T(kMinCosTheta)) return T(0);
        if (alphaT == T(0)) return T(0);
        bool isR = wo.z > T(0);
        V3<T> h = bsdfHalf(wi, wo, isR);
        T wiDotH = dot(wi, h), woDotH = dot(wo, h);
        T cosT;
        T F = fresnelDielectric(eta, wiDotH, cosT);
        T pdf = evalPdfGGX_VNDF(alphaT, wi, h);
        if (isR) {
            pdf /= T(4) * woDotH;
        } else {
            if (woDotH > T(0)) return T(0);
            T sqrtDenom = woDotH + eta * wiDotH;
            T denom = sqrtDenom * sqrtDenom;
            pdf = (denom > T(0)) ? pdf * std::fabs(woDotH) / denom : T(std::numeric_limits<float>::max());
        }
        return pdf * (isR ? F : T(1) - F);
    }
    // StandardBSDF.slang:228-240
    V3<T> eval(V3<T> wi, V3<T> wo) const {
        V3<T> r(T(0));
        if (pDiff > T(0)) r = r + evalDiffuse(wi, wo) * ((T(1) - specTrans) * (T(1) - diffTrans));
        if (pDiffT > T(0)) r = r + evalDiffuseT(wi, wo) * ((T(1) - specTrans) * diffTrans);
        if (pSpec > T(0)) r = r + evalSpecular(wi, wo) * (T(1) - specTrans);
        if (pSpecT > T(0)) r = r + evalSpecularT(wi, wo) * specTrans;
        return r;
    }
    T evalPdf(V3<T> wi, V3<T> wo) const {
        T p = T(0);
        if (pDiff > T(0)) p += pDiff * pdfDiffuse(wi, wo);
        if (pDiffT > T(0)) p += pDiffT * pdfDiffuseT(wi, wo);
        if (pSpec > T(0)) p += pSpec * pdfSpecular(wi, wo);
        if (pSpecT > T(0)) p += pSpecT * pdfSpecularT(wi, wo);
        return p;
    }
    // StandardBSDF.slang:242-313 (one-sample lobe selection).
    bool sample(V3<T> wi, V3<T>& wo, T& pdf, V3<T>& weight, bool& delta, SampleGen& sg) const {
        wo = V3<T>(T(0)); weight = V3<T>(T(0)); pdf = T(0); delta = false;
        bool valid = false;
        T uSelect = T(sg.next1D());
        if (uSelect < pDiff) {
            // FrostbiteDiffuseBRDF.slang:53-72
            T ux = T(sg.next1D()), uy = T(sg.next1D());
            wo = sample_cosine_hemisphere_concentric(ux, uy, pdf);
            if (std::min(wi.z, wo.z) < T(kMinCosTheta)) { weight = V3<T>(T(0)); valid = false; }
            else { weight = frostbiteWeight(wi, wo); valid = true; }
            weight = weight / pDiff;
            weight = weight * ((T(1) - specTrans) * (T(1) - diffTrans));
            pdf *= pDiff;
            if (pSpec > T(0)) pdf += pSpec * pdfSpecular(wi, wo);
            if (pSpecT > T(0)) pdf += pSpecT * pdfSpecularT(wi, wo);
        } else if (uSelect < pDiff + pDiffT) {
            // LambertDiffuseBTDF.slang:49-69
            T ux = T(sg.next1D()), uy = T(sg.next1D());
            wo = sample_cosine_hemisphere_concentric(ux, uy, pdf);
            wo.z = -wo.z;
            if (std::min(wi.z, -wo.z) < T(kMinCosTheta)) { weight = V3<T>(T(0)); valid = false; }
            else { weight = transAlbedo; valid = true; }
            weight = weight / pDiffT;
            weight = weight * ((T(1) - specTrans) * diffTrans);
            pdf *= pDiffT;
            if (pSpecT > T(0)) pdf += pSpecT * pdfSpecularT(wi, wo);
        } else if (uSelect < pDiff + pDiffT + pSpec) {
            valid = sampleSpecular(wi, wo, pdf, weight, delta, sg);
            weight = weight / pSpec;
            weight = weight * (T(1) - specTrans);
            pdf *= pSpec;
            if (pDiff > T(0)) pdf += pDiff * pdfDiffuse(wi, wo);
            if (pSpecT > T(0)) pdf += pSpecT * pdfSpecularT(wi, wo);
        } else if (pSpecT > T(0)) {
            valid = sampleSpecularT(wi, wo, pdf, weight, delta, sg);
            weight = weight / pSpecT;
            weight = weight * specTrans;
            pdf *= pSpecT;
            if (pDiff > T(0)) pdf += pDiff * pdfDiffuse(wi, wo);
            if (pDiffT > T(0)) pdf += pDiffT * pdfDiffuseT(wi, wo);
            if (pSpec > T(0)) pdf += pSpec * pdfSpecular(wi, wo);
        }
        return valid;
    }
    // SpecularMicrofacet.slang:260-410 with both reflection and transmission active: the Fresnel
    // term picks the event; delta events when alpha is 0.
    bool sampleSpecularT(V3<T> wi, V3<T>& wo, T& pdf, V3<T>& weight, bool& delta, SampleGen& sg) const {
        wo = V3<T>(T(0)); weight = V3<T>(T(0)); pdf = T(0);
        if (wi.z < T(kMinCosTheta)) return false;
        T lobeSample = T(sg.next1D());
        if (alphaT == T(0)) {
            T cosT;
            T F = fresnelDielectric(eta, wi.z, cosT);
            bool isR = lobeSample < F;
            pdf = T(0);
            weight = isR ? V3<T>(T(1)) : transAlbedo;
            wo = isR ? V3<T>(-wi.x, -wi.y, wi.z) : V3<T>(-wi.x * eta, -wi.y * eta, -cosT);
            delta = true;
            return !(std::fabs(wo.z) < T(kMinCosTheta) || ((wo.z > T(0)) != isR));
        }
        T ux = T(sg.next1D()), uy = T(sg.next1D());
        V3<T> h = sampleGGX_VNDF(alphaT, wi, ux, uy, pdf);
        T wiDotH = dot(wi, h);
        T cosT;
        T F = fresnelDielectric(eta, wiDotH, cosT);
        bool isR = lobeSample < F;
        wo = isR ? (T(2) * wiDotH * h - wi) : ((eta * wiDotH - cosT) * h - eta * wi);
        if (std::fabs(wo.z) < T(kMinCosTheta) || ((wo.z > T(0)) != isR)) return false;
        T woDotH = dot(wo, h);
        T G = evalMaskingSmithGGXCorrelated(alphaT, wi.z, std::fabs(wo.z));
        T GOverG1wo = G * (T(1) + evalLambdaGGX(alphaT * alphaT, wi.z));
        weight = V3<T>(GOverG1wo);
        if (isR) {
            pdf /= T(4) * woDotH;
        } else {
            T sqrtDenom = woDotH + eta * wiDotH;
            T denom = sqrtDenom * sqrtDenom;
            pdf = (denom > T(0)) ? pdf * std::fabs(woDotH) / denom : T(std::numeric_limits<float>::max());
            weight = weight * (transAlbedo * eta * eta);
        }
        pdf *= isR ? F : T(1) - F;
        return true;
    }
    // SpecularMicrofacet.slang:85-157 (VNDF sampling; delta reflection when alpha == 0).
    bool sampleSpecular(V3<T> wi, V3<T>& wo, T& pdf, V3<T>& weight, bool& delta, SampleGen& sg) const {
        wo = V3<T>(T(0)); weight = V3<T>(T(0)); pdf = T(0);
        if (wi.z < T(kMinCosTheta)) return false;
        if (alpha == T(0)) {
            wo = V3<T>(-wi.x, -wi.y, wi.z);
            pdf = T(0);
            weight = fresnelSchlick3(specAlbedo, T(1), wi.z);
            delta = true;
            return true;
        }
        T ux = T(sg.next1D()), uy = T(sg.next1D());
        V3<T> h = sampleGGX_VNDF(alpha, wi, ux, uy, pdf);
        T wiDotH = dot(wi, h);
        wo = T(2) * wiDotH * h - wi;
        if (wo.z < T(kMinCosTheta)) return false;
        T G = evalMaskingSmithGGXCorrelated(alpha, wi.z, wo.z);
        T GOverG1wo = G * (T(1) + evalLambdaGGX(alpha * alpha, wi.z));
        V3<T> F = fresnelSchlick3(specAlbedo, T(1), wiDotH);
        pdf /= (T(4) * wiDotH);
        weight = F * GOverG1wo;
        return true;
    }
};

// ---------------------------------------------------------------- shading frame
template <typename T> struct Frame {
    V3<T> T_, B, N;
    V3<T> toLocal(V3<T> v) const { return V3<T>(dot(v, T_), dot(v, B), dot(v, N)); }
    V3<T> fromLocal(V3<T> v) const { return T_ * v.x + B * v.y + N * v.z; }
};
// ShadingFrame.slang:76-82 + :127-156 with an invalid tangent (scenes carry no tangents):
// T = perp_stark(N), B = cross(N, T).
template <typename T> inline Frame<T> frameFromNormal(V3<T> n) {
    Frame<T> f;
    f.N = normalize(n);
    f.T_ = perp_stark(f.N);
    f.B = cross(f.N, f.T_);
    return f;
}

}  // namespace fo
