// WARDiffPathTracer primal and adjoint paths (one path per thread; per-path code shared by the
// kernels and the host check build).
//
// Primal: WARDiffPathTracer.rt.slang:75-143 with PTUtils.slang:97-341 (uniform emissive NEE,
// MIS, BSDF sampling, Gaussian pixel filter with antithetic jitter), useWAR = false.
//
// Adjoint: path replay backpropagation. Each thread first traces its path to get the path
// radiance L, then replays it from the same seed. Material parameters enter the path only
// through BSDF evaluations f and emission E, with sampling, pdfs and MIS weights detached as in
// the reference (PTUtils.slang:136-249, StandardBSDF.slang:155-178 `detach`). With
// S = dLdI . L and R_j = radiance gathered after vertex j,
//   dS/df_j(w_j) = dLdI * R_j / f_j(w_j)          (scatter throughput factor)
//   dS/df_j(w_l) = dLdI * thp_j * Li * mis        (NEE)
//   dS/dE_j      = dLdI * thp_j * mis             (emission)
// R_j is obtained as L - (radiance accumulated up to vertex j), so no vertex is stored. The
// first pass records each traversal's answer in a per-thread hit log (4 bytes per path vertex),
// and the replay reads it instead of tracing again: the replay recomputes every ray bit for bit,
// so the answers are the ones a second traversal would give.
// Parameter gradients are routed like GradientIOWrapper.slang:48-116: 20 floats per material
// (StandardMaterialParamLayout.slang:32-41), hash row jenkinsHash(pixelID) % hashSize; the
// per-texel baseColor gradient (texture extension) uses the bilinear tap weights.
#pragma once
#include "fhp_internal.h"
#include "fhp_trace.h"

namespace fhp {
namespace {


// Traversal answers of one path, written by the adjoint's first pass and read by its replay.
// Entry k (at p[k * stride]) is the triangle of closest-hit trace k (trace 0 is the primary ray,
// trace k the scatter ray of vertex k - 1), or kLogMiss; bit 31 is the visibility of vertex k's
// light sample. Entries at k >= depth are not kept: those rays are traced in both passes.
constexpr uint32_t kLogMiss = 0x7fffffffu;
constexpr uint32_t kLogLit = 0x80000000u;
struct HitLog {
    GlobalU32* p;     // this thread's entry 0, or null (both passes trace every ray)
    uint32_t stride;  // threads of the persistent grid
    uint32_t depth;   // entries per thread
    FHP_DEVI bool has(uint32_t k) const { return p != nullptr && k < depth; }
    FHP_DEVI void put(uint32_t k, uint32_t v) const { p[k * stride] = v; }
    FHP_DEVI uint32_t get(uint32_t k) const { return p[k * stride]; }
};
FHP_DEVI HitLog noLog() { return HitLog{nullptr, 0u, 0u}; }

// Gaussian pixel jitter (WARDiffPathTracer.rt.slang:60-63).
FHP_DEVI void gaussian2(float ux, float uy, float sigma, float& gx, float& gy) {
    float r = sqrtf(-2.f * logf(ux));
    gx = (r * cosf(kTwoPi * uy)) * sigma;
    gy = (r * sinf(kTwoPi * uy)) * sigma;
}

// Local gradient of one vertex's material parameters. ior holds dS/dF0 (converted to dS/dior when
// the vertex is flushed); iorEta is dS/dior through the relative IoR of the transmission lobes.
struct MatGrad {
    f3 base;
    float metal, rough, ior;
    f3 emit;
    float emitFactor;
    f3 trans;                          // transmissionColor
    float diffTrans, specTrans, iorEta;
};
FHP_DEVI void zero(MatGrad& g) {
    g.base = mk3(0.f); g.metal = g.rough = g.ior = 0.f; g.emit = mk3(0.f); g.emitFactor = 0.f;
    g.trans = mk3(0.f); g.diffTrans = g.specTrans = g.iorEta = 0.f;
}

// ------------------------------------------------------------------ transmission lobes
// Forward-mode dual numbers over (alpha, eta): value and the two partial derivatives.
struct D2 {
    float v, a, e;
};
FHP_DEVI D2 d2c(float c) { return D2{c, 0.f, 0.f}; }
FHP_DEVI D2 operator+(D2 x, D2 y) { return D2{x.v + y.v, x.a + y.a, x.e + y.e}; }
FHP_DEVI D2 operator-(D2 x, D2 y) { return D2{x.v - y.v, x.a - y.a, x.e - y.e}; }
FHP_DEVI D2 operator-(D2 x) { return D2{-x.v, -x.a, -x.e}; }
FHP_DEVI D2 operator*(D2 x, D2 y) { return D2{x.v * y.v, x.a * y.v + x.v * y.a, x.e * y.v + x.v * y.e}; }
FHP_DEVI D2 operator*(D2 x, float c) { return D2{x.v * c, x.a * c, x.e * c}; }
FHP_DEVI D2 operator*(float c, D2 x) { return x * c; }
FHP_DEVI D2 operator+(D2 x, float c) { return D2{x.v + c, x.a, x.e}; }
FHP_DEVI D2 operator+(float c, D2 x) { return x + c; }
FHP_DEVI D2 operator-(float c, D2 x) { return D2{c - x.v, -x.a, -x.e}; }
FHP_DEVI D2 operator/(D2 x, D2 y) {
    const float q = x.v / y.v;
    return D2{q, (x.a - q * y.a) / y.v, (x.e - q * y.e) / y.v};
}
FHP_DEVI D2 d2sqrt(D2 x) {
    const float r = sqrtf(x.v), k = r > 0.f ? 0.5f / r : 0.f;
    return D2{r, x.a * k, x.e * k};
}
FHP_DEVI D2 d2abs(D2 x) { return x.v < 0.f ? -x : x; }
// IsotropicGGX.slang:46-51, :163-169, :200-206 with alpha a dual number (cosines ci, co fixed).
FHP_DEVI D2 ggxDd(D2 alpha, D2 c) {
    const D2 a2 = alpha * alpha;
    const D2 d = (c * a2 - c) * c + 1.f;
    return a2 / (d * d * kPi);
}
FHP_DEVI D2 ggxLambdad(D2 a2, float c) {
    const float c2 = c * c;
    const float t2 = fmaxf_(1.f - c2, 0.f) / c2;
    return (-1.f + d2sqrt(1.f + a2 * t2)) * 0.5f;
}
// Fresnel.slang:68-92 with eta and cosI dual numbers (total internal reflection: a constant 1).
FHP_DEVI D2 fresnelDielectricd(D2 eta, D2 cosI) {
    if (cosI.v < 0.f) {
        eta = d2c(1.f) / eta;
        cosI = -cosI;
    }
    const D2 sin2T = eta * eta * (1.f - cosI * cosI);
    if (sin2T.v > 1.f) return d2c(1.f);
    const D2 cosT = d2sqrt(1.f - sin2T);
    const D2 rs = (eta * cosI - cosT) / (eta * cosI + cosT);
    const D2 rp = (eta * cosT - cosI) / (eta * cosT + cosI);
    return 0.5f * (rs * rs + rp * rp);
}
// The GGX dielectric lobe of Bsdf::evalSpecularT (SpecularMicrofacetBSDF::eval) as a dual number
// in (alphaT, eta), without the transmission albedo: on the reflection side F D G / (4 wi.z), on
// the transmission side (1 - F) D G |t| (times the albedo per channel). Same operations as the
// float evaluation; the caller has checked the lobe's cosine and alpha conditions.
FHP_DEVI D2 specTDual(f3 wi, f3 wo, float alphaT, float eta, bool& isR) {
    isR = wo.z > 0.f;
    const D2 A{alphaT, 1.f, 0.f}, E{eta, 0.f, 1.f};
    const D2 k = isR ? d2c(1.f) : E;
    D2 hx = d2c(wo.x) + k * wi.x, hy = d2c(wo.y) + k * wi.y, hz = d2c(wo.z) + k * wi.z;
    const D2 len = d2sqrt(hx * hx + hy * hy + hz * hz);
    hx = hx / len;
    hy = hy / len;
    hz = hz / len;
    const float sgn = hz.v > 0.f ? 1.f : (hz.v < 0.f ? -1.f : 0.f);
    hx = hx * sgn;
    hy = hy * sgn;
    hz = hz * sgn;
    const D2 wiDotH = hx * wi.x + hy * wi.y + hz * wi.z, woDotH = hx * wo.x + hy * wo.y + hz * wo.z;
    const D2 D = ggxDd(A, hz);
    const D2 a2 = A * A;
    const D2 G = d2c(1.f) / (1.f + ggxLambdad(a2, wi.z) + ggxLambdad(a2, fabsf(wo.z)));
    const D2 F = fresnelDielectricd(E, wiDotH);
    if (isR) return F * D * G * (0.25f / wi.z);
    const D2 sqrtDenom = woDotH + E * wiDotH;
    const D2 t = E * E * wiDotH * woDotH / (sqrtDenom * sqrtDenom * wi.z);
    return (1.f - F) * D * G * d2abs(t);
}

// Adjoint of the four-lobe StandardBSDF (StandardBSDF.slang:125-240) of a material with
// transmission, f = (1 - sT)(1 - dT) f_dr + (1 - sT) dT f_dt + (1 - sT) f_sr + sT f_st with the
// lobes gated by their (detached) selection probabilities, into baseColor, metallic, roughness,
// IoR (through F0 and through the relative IoR of f_st), transmissionColor (the albedo is its
// square root unless the surface is thin), diffuseTransmission and specularTransmission.
FHP_DEV void bsdfAdjointTrans(const Shading& sd, const Material& m, bool thin, f3 wo, f3 adj, MatGrad& g) {
    f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(wo);
    if (!validRT(sd, wiL, woL, wo)) return;
    Bsdf b;
    b.init(wiL, m);
    const float sT = b.specTrans, dT = b.diffTrans;
    const float wDR = (1.f - sT) * (1.f - dT), wDT = (1.f - sT) * dT, wSR = 1.f - sT, wST = sT;
    f3 fDR = mk3(0.f), fDT = mk3(0.f), fSR = mk3(0.f), fST = mk3(0.f);
    f3 dDiff = mk3(0.f), dSpec = mk3(0.f), dTA = mk3(0.f);
    float dRough = 0.f, dEta = 0.f;
    if (b.pDiff > 0.f && fminf_(wiL.z, woL.z) >= kMinCosTheta) {
        // Frostbite: f = diff * wiS * woS * eF / pi * wo.z
        f3 h = normalize(wiL + woL);
        float woDotH = dot(woL, h);
        float r = b.rough;
        float eF = 1.f + (1.f / 1.51f - 1.f) * r;
        float fd90 = 0.5f * r + 2.f * woDotH * woDotH * r;
        float wi5 = pow5(fmaxf_(1.f - wiL.z, 0.f)), wo5 = pow5(fmaxf_(1.f - woL.z, 0.f));
        float wiS = 1.f + (fd90 - 1.f) * wi5, woS = 1.f + (fd90 - 1.f) * wo5;
        float c = kInvPi * woL.z;
        float K = wiS * woS * eF * c;
        fDR = b.diff * K;
        dDiff = adj * (K * wDR);
        float dfd90 = 0.5f + 2.f * woDotH * woDotH;
        float dProd = (dfd90 * wi5) * woS * eF + wiS * (dfd90 * wo5) * eF + wiS * woS * (1.f / 1.51f - 1.f);
        dRough += dot(adj, b.diff) * dProd * c * wDR;
    }
    if (b.pDiffT > 0.f && fminf_(wiL.z, -woL.z) >= kMinCosTheta) {
        // LambertDiffuseBTDF: f = albedo / pi * |wo.z|
        const float k = kInvPi * -woL.z;
        fDT = b.trans * k;
        dTA = dTA + adj * (k * wDT);
    }
    if (b.pSpec > 0.f && b.alpha > 0.f && fminf_(wiL.z, woL.z) >= kMinCosTheta) {
        f3 h = normalize(wiL + woL);
        float wiDotH = dot(wiL, h);
        float D = ggxD(b.alpha, h.z), G = ggxG(b.alpha, wiL.z, woL.z);
        float w5 = pow5(fmaxf_(1.f - wiDotH, 0.f));
        float sc = D * G * 0.25f / wiL.z;
        f3 F = schlick3(b.spec, 1.f, wiDotH);
        fSR = F * sc;
        dSpec = adj * ((1.f - w5) * sc * wSR);
        float a2 = b.alpha * b.alpha, cz = h.z;
        float d = (cz * a2 - cz) * cz + 1.f;
        float dD = (d - 2.f * a2 * cz * cz) / (kPi * d * d * d);
        auto dLam = [&](float c0) {
            float c2 = c0 * c0;
            float t2 = fmaxf_(1.f - c2, 0.f) / c2;
            return t2 / (4.f * sqrtf(1.f + a2 * t2));
        };
        float dG = -G * G * (dLam(wiL.z) + dLam(woL.z));
        float dA2 = (dD * G + D * dG) * 0.25f / wiL.z;
        float r = b.rough;
        dRough += dot(adj, F) * dA2 * (4.f * r * r * r) * wSR;
    }
    if (b.pSpecT > 0.f && b.alphaT > 0.f && fminf_(wiL.z, fabsf(woL.z)) >= kMinCosTheta) {
        bool isR;
        const D2 v = specTDual(wiL, woL, b.alphaT, b.eta, isR);
        const f3 col = isR ? mk3(1.f) : b.trans;
        fST = col * v.v;
        const float ac = dot(adj, col) * wST;
        dRough += ac * v.a * (2.f * b.rough);  // alphaT = roughness^2 (eta != 1, alpha >= the delta bound)
        dEta += ac * v.e;
        if (!isR) dTA = dTA + adj * (v.v * wST);
    }
    // Mixing weights (the selection probabilities that gate the lobes are detached).
    g.diffTrans += dot(adj, (fDT - fDR) * (1.f - sT));
    g.specTrans += dot(adj, fST - fDR * (1.f - dT) - fDT * dT - fSR);
    // diffuse = base (1 - metal); specular = F0 + (base - F0) metal.
    g.base = g.base + dDiff * (1.f - m.metallic) + dSpec * m.metallic;
    g.metal += -dot(dDiff, m.base) + dot(dSpec, m.base - mk3(m.F0));
    g.rough += dRough;
    g.ior += (dSpec.x + dSpec.y + dSpec.z) * (1.f - m.metallic);
    // eta = 1 / IoR on the front side, IoR on the back side (StandardMaterial.slang:147).
    g.iorEta += sd.frontFacing ? -dEta * b.eta * b.eta : dEta;
    // Transmission albedo: sqrt(colour), or the colour on thin surfaces (StandardBSDF.slang:131).
    const f3 tA = b.trans;
    if (thin) g.trans = g.trans + dTA;
    else g.trans = g.trans + mk3(tA.x > 0.f ? dTA.x * 0.5f / tA.x : 0.f, tA.y > 0.f ? dTA.y * 0.5f / tA.y : 0.f,
                                 tA.z > 0.f ? dTA.z * 0.5f / tA.z : 0.f);
}

// Backpropagates adj (dS/df, RGB) through f = material BSDF eval(wi, wo) into material params
// (reflection lobes; bsdfAdjointTrans above handles materials with transmission). The mixing
// weights are differentiated at specularTransmission = diffuseTransmission = 0 too, as the
// reference reads both through its gradient wrapper for every StandardMaterial
// (StandardMaterial.slang:148-149): df/d(diffTrans) = -f_dr, df/d(specTrans) = -(f_dr + f_sr).
FHP_DEV void bsdfAdjoint(const Shading& sd, const Material& m, f3 wo, f3 adj, MatGrad& g) {
    f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(wo);
    if (m.type == MAT_PBRT_DIFFUSE) {
        if (!validR(sd, wiL, woL, wo)) return;
        float k = fmaxf_(woL.z, 0.f) * kInvPi;
        g.base = g.base + adj * k;
        return;
    }
    if (!validRT(sd, wiL, woL, wo)) return;
    Bsdf b;
    b.init(wiL, m);
    f3 dDiff = mk3(0.f), dSpec = mk3(0.f);
    float dRough = 0.f, fDR = 0.f, fSR = 0.f;  // adj . f_dr, adj . f_sr
    if (b.pDiff > 0.f && fminf_(wiL.z, woL.z) >= kMinCosTheta) {
        // Frostbite: f = diff * wiS * woS * eF / pi * wo.z
        f3 h = normalize(wiL + woL);
        float woDotH = dot(woL, h);
        float r = b.rough;
        float eF = 1.f + (1.f / 1.51f - 1.f) * r;
        float fd90 = 0.5f * r + 2.f * woDotH * woDotH * r;
        float wi5 = pow5(fmaxf_(1.f - wiL.z, 0.f)), wo5 = pow5(fmaxf_(1.f - woL.z, 0.f));
        float wiS = 1.f + (fd90 - 1.f) * wi5, woS = 1.f + (fd90 - 1.f) * wo5;
        float c = kInvPi * woL.z;
        float K = wiS * woS * eF * c;
        dDiff = adj * K;
        fDR = dot(adj, b.diff) * K;
        float dfd90 = 0.5f + 2.f * woDotH * woDotH;
        float dProd = (dfd90 * wi5) * woS * eF + wiS * (dfd90 * wo5) * eF + wiS * woS * (1.f / 1.51f - 1.f);
        dRough += dot(adj, b.diff) * dProd * c;
    }
    if (b.pSpec > 0.f && b.alpha > 0.f && fminf_(wiL.z, woL.z) >= kMinCosTheta) {
        f3 h = normalize(wiL + woL);
        float wiDotH = dot(wiL, h);
        float D = ggxD(b.alpha, h.z), G = ggxG(b.alpha, wiL.z, woL.z);
        float w5 = pow5(fmaxf_(1.f - wiDotH, 0.f));
        float s = D * G * 0.25f / wiL.z;
        dSpec = adj * ((1.f - w5) * s);
        // d/d(alpha^2) of D and G.
        float a2 = b.alpha * b.alpha, cz = h.z;
        float d = (cz * a2 - cz) * cz + 1.f;
        float dD = (d - 2.f * a2 * cz * cz) / (kPi * d * d * d);
        auto dLam = [&](float c0) {
            float c2 = c0 * c0;
            float t2 = fmaxf_(1.f - c2, 0.f) / c2;
            return t2 / (4.f * sqrtf(1.f + a2 * t2));
        };
        float dG = -G * G * (dLam(wiL.z) + dLam(woL.z));
        f3 F = schlick3(b.spec, 1.f, wiDotH);
        float dA2 = (dD * G + D * dG) * 0.25f / wiL.z;
        float r = b.rough;
        dRough += dot(adj, F) * dA2 * (4.f * r * r * r);
        fSR = dot(adj, F) * s;
    }
    // diffuse = base (1 - metal); specular = F0 + (base - F0) metal.
    g.base = g.base + dDiff * (1.f - m.metallic) + dSpec * m.metallic;
    g.metal += -dot(dDiff, m.base) + dot(dSpec, m.base - mk3(m.F0));
    g.rough += dRough;
    // dS/dF0, converted to dS/dior = dS/dF0 * dF0/dior when the vertex is flushed.
    g.ior += (dSpec.x + dSpec.y + dSpec.z) * (1.f - m.metallic);
    g.diffTrans -= fDR;
    g.specTrans -= fDR + fSR;
}

FHP_DEVI void atomAdd(float* p, float v) {
    if (v != 0.f) gradAdd(p, v);
}

// BSDF adjoint of a vertex: the transmission lobes' adjoint for a StandardMaterial with
// transmission (compiled in the TRANS kernels only), the reflection lobes' otherwise.
template <bool TRANS>
FHP_DEVI void matAdjoint(const SceneDev& s, const Shading& sd, const Material& m, f3 wo, f3 adj, MatGrad& g) {
    if (TRANS && m.type == MAT_STANDARD && (m.specTrans > 0.f || m.diffTrans > 0.f))
        bsdfAdjointTrans(sd, m, s.mats[size_t(sd.mat) * M_STRIDE + M_THIN] != 0.f, wo, adj, g);
    else
        bsdfAdjoint(sd, m, wo, adj, g);
}

// The gradient-image projection of one vertex's gradient (the debug modes): the entry of the
// 20-float layout that starts at a.dbgOffset of material a.dbgMat, dotted with a.dbgGrad, as
// GradientIOWrapper seeds or reads it (GradientIOWrapper.slang:75-106); zero for other materials
// and for offsets that start no entry.
FHP_DEVI float debugProjection(const SceneDev& s, const WarArgs& a, const Material& m, uint32_t mat, const MatGrad& g) {
    if (mat != a.dbgMat) return 0.f;
    const float* v = a.dbgGrad;
    switch (a.dbgOffset) {
        case 0: return g.base.x * v[0] + g.base.y * v[1] + g.base.z * v[2];
        default: break;
    }
    if (m.type != MAT_STANDARD) return 0.f;
    switch (a.dbgOffset) {
        case 3: return g.metal * v[0];
        case 4: return g.rough * v[0];
        case 5: {
            const float ior = s.mats[size_t(mat) * M_STRIDE + M_IOR];
            const float f = (ior - 1.f) / (ior + 1.f);
            return (g.ior * 2.f * f * (2.f / ((ior + 1.f) * (ior + 1.f))) + g.iorEta) * v[0];
        }
        case 6: return g.trans.x * v[0] + g.trans.y * v[1] + g.trans.z * v[2];
        case 9: return g.diffTrans * v[0];
        case 10: return g.specTrans * v[0];
        case 11: return g.emit.x * v[0] + g.emit.y * v[1] + g.emit.z * v[2];
        case 14: return g.emitFactor * v[0];
        default: return 0.f;
    }
}

FHP_DEV void flushGrad(const SceneDev& s, const WarArgs& a, uint32_t hashRow, const Material& m, uint32_t mat,
                          const TexTap& tap, const MatGrad& g) {
    float* row = a.gradMat + size_t(hashRow) * a.matGradDim + size_t(mat) * 20;
    atomAdd(row + 0, g.base.x);
    atomAdd(row + 1, g.base.y);
    atomAdd(row + 2, g.base.z);
    if (m.type == MAT_STANDARD) {
        atomAdd(row + 3, g.metal);
        atomAdd(row + 4, g.rough);
        const float ior = s.mats[size_t(mat) * M_STRIDE + M_IOR];
        const float f = (ior - 1.f) / (ior + 1.f);
        const float dfdior = 2.f / ((ior + 1.f) * (ior + 1.f));
        atomAdd(row + 5, g.ior * 2.f * f * dfdior + g.iorEta);
        atomAdd(row + 6, g.trans.x);
        atomAdd(row + 7, g.trans.y);
        atomAdd(row + 8, g.trans.z);
        atomAdd(row + 9, g.diffTrans);
        atomAdd(row + 10, g.specTrans);
        atomAdd(row + 11, g.emit.x);
        atomAdd(row + 12, g.emit.y);
        atomAdd(row + 13, g.emit.z);
        atomAdd(row + 14, g.emitFactor);
    }
    if (a.gradTex && m.tex >= 0) {
        float* t = a.gradTex + tap.base;
        for (int k = 0; k < 4; k++) {
            float w = tap.w[k];
            if (w == 0.f) continue;
            float* q = t + 3 * size_t(tap.t[k]);
            atomAdd(q + 0, g.base.x * w);
            atomAdd(q + 1, g.base.y * w);
            atomAdd(q + 2, g.base.z * w);
        }
    }
}

// One path of WARDiffPathTracer (tracePaths + tracePath). In ADJ mode, Ltot is the path radiance
// from a previous primal run and gradients are accumulated; returns the path radiance.
//
// Each vertex first computes the differentiable intersection, draws the light sample and traces
// its visibility ray, and only then loads the shading data and material and evaluates emission,
// NEE and the scatter sample: the heavy per-vertex state (shading frame, material, gradient
// accumulator, texture taps) is never live across a traversal, which keeps the kernels within
// their register budget. The random numbers are drawn and the radiance terms summed in the
// reference's order (emission, NEE, scatter), so the results do not depend on this order of
// evaluation. FROZEN: the finite-difference mode's separate sampling parameters (matsS /
// texelsS) are in use; otherwise one material serves evaluation and sampling.
// DEBUG (with ADJ): instead of accumulating the gradients, each vertex adds its debug projection
// to *dproj. With a hit log, the pass without ADJ records its traversal answers and the ADJ pass
// replays them (tracing only the rays beyond the log's depth).
template <bool ADJ, bool FROZEN, bool DEBUG = false, bool TRANS = false>
FHP_DEV f3 warPath(const SceneDev& s, const WarArgs& a, uint32_t view, uint32_t px, uint32_t py,
                   uint32_t frameSeed, TraceStack& st, f3 Ltot, f3 dL, uint32_t hashRow, unsigned long long& nc,
                   unsigned long long& nv, const HitLog& hl, float* dproj = nullptr) {
    Rng rng = rngSeed(px, py, frameSeed / 2);
    for (int i = 0; i < 6; i++) rng.u1();
    const float pfx = float(px) + 0.5f, pfy = float(py) + 0.5f;
    float ux = rng.u1(), uy = rng.u1();
    float jx, jy;
    gaussian2(ux, uy, 0.5f, jx, jy);
    if (a.useAntithetic && frameSeed % 2 == 0) { jx = -jx; jy = -jy; }
    const float sx = pfx + jx, sy = pfy + jy;
    f3 radiance = mk3(0.f), thp = mk3(1.f);
    uint32_t len = 0;
    f3 prevN = mk3(0.f);
    float prevPdf = 0.f;
    // computeCameraRayDirection (DiffSceneQuery.slang:252-261), with the view's camera.
    const CameraDev cam = a.views ? a.views[view].cam : s.cam;
    const float p_x = sx / float(a.W), p_y = sy / float(a.H);
    const float ndc_x = 2.f * p_x + -1.f, ndc_y = -2.f * p_y + 1.f;
    f3 dir = cam.U * ndc_x + cam.V * ndc_y + cam.W;
    const float dl = length(dir);
    dir = dl > 0.f ? dir / dl : mk3(0.f);
    f3 rayO = cam.pos, rayD = dir;
    HitRec h;
    // Closest-hit trace k: from the log in the replay, recorded in the first pass.
    auto closest = [&](uint32_t k) -> bool {
        if (ADJ && hl.has(k)) {
            const uint32_t e = hl.get(k) & ~kLogLit;
            h.tri = e;
            return e != kLogMiss;
        }
        nc++;
        const bool hit = traceBVH4<false, false>(s, rayO, rayD, 0.f, kFltMax, st, h);
        if (!ADJ && hl.has(k) && !hit) hl.put(k, kLogMiss);
        return hit;
    };
    if (!closest(0)) return radiance;
    for (;;) {
        // computeIntersectionAD, AttachToRay (DiffSceneQuery.slang:131-219).
        f3 posW, normalW, bary;
        {
            const uint32_t i0 = s.idx[3 * h.tri], i1 = s.idx[3 * h.tri + 1], i2 = s.idx[3 * h.tri + 2];
            f3 v0, v1, v2, n0, n1, n2;
            float tu, tv;
            loadVertex(s, i0, v0, n0, tu, tv);
            loadVertex(s, i1, v1, n1, tu, tv);
            loadVertex(s, i2, v2, n2, tu, tv);
            f3 e1 = v1 - v0, e2 = v2 - v0;
            f3 pvec = cross(rayD, e2);
            float det = dot(e1, pvec);
            f3 tvec = rayO - v0;
            f3 qvec = cross(tvec, e1);
            float tt = dot(e2, qvec) / det;
            f3 p = rayO + rayD * tt;
            f3 ep0 = v0 - p, ep1 = v1 - p, ep2 = v2 - p;
            float A0 = length(cross(ep1, ep2)), A1 = length(cross(ep2, ep0)), A2 = length(cross(ep0, ep1));
            float At = A0 + A1 + A2;
            bary = mk3(A0 / At, A1 / At, A2 / At);
            posW = v0 * bary.x + v1 * bary.y + v2 * bary.z;
            normalW = n0 * bary.x + n1 * bary.y + n2 * bary.z;
        }
        const bool stop = len >= a.maxBounces + 1;
        // generateEmissiveSample + traceVisibilityRay (PTUtils.slang:177-207, :310-319).
        LightSample ls;
        bool lit = false;
        f3 Li = mk3(0.f), ldir = mk3(0.f);
        if (!stop && a.useNEE && sampleLight(s, LS_UNIFORM, posW, normalW, rng, ls)) {
            Li = ls.pdf > 0.f ? ls.Le / ls.pdf : mk3(0.f);
            const f3 lightPos = rayOriginOffset(ls.posW, ls.normalW);
            const f3 origin = rayOriginOffset(posW, normalW);
            const f3 toL = lightPos - origin;
            const float distance = length(toL);
            ldir = normalize(toL);
            if (anyPos(Li)) {
                if (ADJ && hl.has(len)) {
                    lit = (hl.get(len) & kLogLit) != 0u;
                } else {
                    HitRec sh;
                    nv++;
                    lit = !traceBVH4<true, false>(s, origin, ldir, 0.f, distance, st, sh);
                }
            }
        }
        if (!ADJ && hl.has(len)) hl.put(len, h.tri | (lit ? kLogLit : 0u));
        // PTUtils.slang:97-123
        Shading sd = shadingFromHit(s, h.tri, bary.y, bary.z, rayD);
        sd.V = -rayD;
        setFrame(sd, normalW);
        TexTap tap;
        Material m, mS;
        {
            const Shading sd0 = sd;
            m = loadMaterial(s, s.mats, s.texels, sd0, sd, ADJ && !DEBUG ? &tap : nullptr);
            // Sampling parameters differ from the evaluated ones only in the finite-difference mode.
            if (FROZEN) {
                Shading unused = sd0;
                mS = loadMaterial(s, s.matsS, s.texelsS, sd0, unused, nullptr);
            }
        }
        const Material& mSamp = FROZEN ? mS : m;
        MatGrad g;
        if (ADJ) zero(g);
        const bool isPrimary = len == 0;
        const bool computeEmissive = isPrimary || (!a.useNEE || a.useMIS);
        if (computeEmissive && anyPos(m.emission)) {
            float mis = 1.f;
            if (!isPrimary && a.useNEE && a.useMIS) {
                float lp = lightPdf(s, LS_UNIFORM, rayO, prevN, s.triToLight[h.tri], sd.posW, sd.orientedFaceN());
                mis = misPower(prevPdf, lp);
            }
            radiance = radiance + thp * m.emission * mis;
            if (ADJ && m.type == MAT_STANDARD && sd.frontFacing) {
                const float* r = s.mats + size_t(sd.mat) * M_STRIDE;
                f3 dE = dL * thp * mis;
                if (r[M_EMIT_TEX] <= 0.f) g.emit = g.emit + dE * r[M_EMIT_FACTOR];  // textured: no colour slot
                g.emitFactor += dot(dE, m.emitColor);
            }
        }
        if (lit) {
            // computeNEE (PTUtils.slang:211-249)
            f3 wo = ls.posW - posW;
            float dist = length(wo);
            wo = wo / dist;
            float a2s = (dist * dist) / dot(ls.normalW, -wo);
            f3 f = FROZEN ? matEvalGated(sd, m, mS, wo) : matEval(sd, m, wo);
            f3 w = Li * f * a2s / a2s;
            float misN = 1.f;
            if (a.useMIS) {
                float sp;
                if (a.useBSDFSampling || m.type == MAT_PBRT_DIFFUSE) sp = matPdf(sd, mSamp, ldir);
                else {
                    f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(ldir);
                    sp = (validRT(sd, wiL, woL, ldir) && fminf_(wiL.z, woL.z) >= kMinCosTheta) ? woL.z * kInvPi : 0.f;
                }
                misN = misPower(ls.pdf, sp);
                w = w * misN;
            }
            radiance = radiance + thp * w;
            if (ADJ) matAdjoint<TRANS>(s, sd, m, wo, dL * thp * Li * (a2s / a2s) * misN, g);
        }
        bool cont = false;
        f3 fS = mk3(0.f), woS = mk3(0.f);
        if (!stop) {
            // generateScatterRay (PTUtils.slang:136-172)
            Scatter sc;
            if (matSample(sd, mSamp, rng, sc, a.useBSDFSampling != 0) && sc.pdf > 0.f) {
                rayD = sc.wo;
                // PTUtils.slang:153: the view side, also for transmitted directions.
                rayO = rayOriginOffset(posW, sd.frontFacing ? sd.faceN : -sd.faceN);
                fS = FROZEN ? matEvalGated(sd, m, mS, rayD) : matEval(sd, m, rayD);
                woS = rayD;
                thp = thp * (fS / sc.pdf);
                prevN = sd.orientedFaceN();
                prevPdf = sc.pdf;
                cont = anyPos(thp);
            }
        }
        if (ADJ) {
            if (cont) {
                f3 R = Ltot - radiance;
                f3 adj = mk3(fS.x != 0.f ? dL.x * R.x / fS.x : 0.f, fS.y != 0.f ? dL.y * R.y / fS.y : 0.f,
                             fS.z != 0.f ? dL.z * R.z / fS.z : 0.f);
                matAdjoint<TRANS>(s, sd, m, woS, adj, g);
            }
            if (DEBUG) *dproj += debugProjection(s, a, m, sd.mat, g);
            else flushGrad(s, a, hashRow, m, sd.mat, tap, g);
        }
        if (!cont) break;
        len++;
        if (!closest(len)) break;
    }
    return radiance;
}

// One item of the kernels' grid loop (path id = pixel * spp + sample): the primal writes the
// path's radiance; the adjoint traces the path, then replays it from the same seed (one launch:
// splitting the two passes into launches measured slower).
// Path id -> (view, path within the view); the view's frame seed base and dL/dI.
struct WarViewRef {
    uint32_t view, frame, local;
    const float* dLdI;
};
template <typename Args> FHP_DEVI WarViewRef warView(const Args& a, uint32_t id) {
    if (!a.views) return WarViewRef{0u, a.frame, id, a.dLdI};
    const uint32_t per = a.W * a.H * a.spp, v = id / per;
    const WarView& w = a.views[v];
    return WarViewRef{v, w.frame, id - v * per, w.dLdI};
}

template <bool ADJ, bool FROZEN, bool TRANS = false>
FHP_DEVI void warItem(const SceneDev& s, const WarArgs& a, uint32_t id, TraceStack& st, unsigned long long& nc,
                      unsigned long long& nv, const HitLog& hl = noLog()) {
    const WarViewRef vr = warView(a, id);
    const uint32_t pix = vr.local / a.spp, k = vr.local % a.spp;
    const uint32_t px = pix % a.W, py = pix / a.W;
    const uint32_t seed = vr.frame + k;
    if (!ADJ) {
        f3 L = warPath<false, FROZEN>(s, a, vr.view, px, py, seed, st, mk3(0.f), mk3(0.f), 0, nc, nv, noLog());
        a.out[id] = make_float4(L.x, L.y, L.z, 0.f);
    } else {
        // dLdI null: the BackwardDiffDebug mode's dL/dI = (1, 0, 0) at every pixel.
        f3 dL = mk3(1.f, 0.f, 0.f);
        if (vr.dLdI) {
            const float* dl = vr.dLdI + 3 * size_t(pix);
            dL = mk3(dl[0], dl[1], dl[2]);
        }
        if (dL.x == 0.f && dL.y == 0.f && dL.z == 0.f) return;
        f3 L = warPath<false, FROZEN>(s, a, vr.view, px, py, seed, st, mk3(0.f), mk3(0.f), 0, nc, nv, hl);
        warPath<true, FROZEN, false, TRANS>(s, a, vr.view, px, py, seed, st, L, dL, jenkins(pix) % a.hashSize, nc, nv,
                                            hl);
    }
}

// One item of the ForwardDiffDebug image kernel: the path's radiance, then its replay with
// dL/dI = (1, 0, 0) projected on the debug direction: the forward-mode derivative of the red
// channel (the reverse-mode gradient of one path estimate, dotted with a direction, is its
// directional derivative). tmp[id] = (radiance, derivative).
template <bool FROZEN, bool TRANS>
FHP_DEVI void warDebugItem(const SceneDev& s, const WarArgs& a, uint32_t id, TraceStack& st, unsigned long long& nc,
                           unsigned long long& nv, const HitLog& hl = noLog()) {
    const uint32_t pix = id / a.spp, k = id % a.spp;
    const uint32_t px = pix % a.W, py = pix / a.W;
    const uint32_t seed = a.frame + k;
    const f3 L = warPath<false, FROZEN>(s, a, 0u, px, py, seed, st, mk3(0.f), mk3(0.f), 0, nc, nv, hl);
    float d = 0.f;
    // The replay's rays are the first pass's: count them once.
    unsigned long long c0 = 0, c1 = 0;
    warPath<true, FROZEN, true, TRANS>(s, a, 0u, px, py, seed, st, L, mk3(1.f, 0.f, 0.f), 0, hl.p ? nc : c0,
                                       hl.p ? nv : c1, hl, &d);
    a.tmp[id] = make_float4(L.x, L.y, L.z, d);
}

// Whether a scene renders with the finite-difference mode's separate sampling parameters.
FHP_HD bool samplingFrozen(const SceneDev& s) { return s.matsS != s.mats || s.texelsS != s.texels; }

}  // namespace
}  // namespace fhp
