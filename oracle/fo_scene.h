// Oracle scene, traversal, emissive-light structures and integrators.
//
// TEST INFRASTRUCTURE ONLY (see fo_core.h). Restates, in scalar C++:
//   * the closest-hit / any-hit queries of the reference's TraceRay / RayQuery
//     (RaytracingInline.slang:86-347) over its own brute-force or median-split BVH2,
//   * shading-data preparation (Scene.slang:440-476, MaterialFactory.slang:78-104),
//   * the LightCollection (LightCollection.cpp:220-262, FinalizeIntegration.cs.slang:103-111),
//   * EmissiveUniformSampler and LightBVHBuilder/LightBVHSampler,
//   * MinimalPathTracer.rt.slang:392-468 and the WARDiffPathTracer primal
//     (WARDiffPathTracer.rt.slang:75-143, PTUtils.slang:253-341).
#pragma once
#include "fo_core.h"
#include <functional>
#include <numeric>

namespace fo {

using V3f = V3<float>;
constexpr uint32_t kInvalid = 0xffffffffu;

// Camera basis exactly as Camera.cpp:187-193 computes it (host FP32 arithmetic).
struct Camera {
    V3f pos, U, V, W;
    float nearZ = 0.1f, farZ = 1000.f;
};
inline Camera makeCamera(V3f pos, V3f target, V3f up, float fovY, float aspect, float focalDistance) {
    Camera c;
    c.pos = pos;
    c.W = normalize(target - pos) * focalDistance;
    c.U = normalize(cross(c.W, up));
    c.V = normalize(cross(c.U, c.W));
    const float ulen = focalDistance * std::tan(fovY * 0.5f) * aspect;
    c.U = c.U * ulen;
    const float vlen = focalDistance * std::tan(fovY * 0.5f);
    c.V = c.V * vlen;
    return c;
}

struct Hit {
    uint32_t tri = kInvalid;
    float t = 0.f, u = 0.f, v = 0.f;  // (u,v) = weights of v1,v2 (HitInfo.slang:76-79)
    bool valid() const { return tri != kInvalid; }
};

struct EmissiveTri {
    V3f p[3];
    float uv[3][2];
    V3f normal;
    float area;
    V3f Le;          // constant emitter: radiance; textured emitter: emissiveFactor (x3)
    int etex = -1;   // emissive texture (MaterialSystem::evalEmissive at the sample's uv), -1 none
    float flux;
    uint32_t tri;
};

struct Scene {
    uint32_t nV = 0, nT = 0, nM = 0, nTex = 0;
    std::vector<V3f> pos, nrm;
    std::vector<float> uv;
    std::vector<uint32_t> idx, triMat;
    std::vector<float> mats, texels;
    std::vector<int32_t> texDesc;
    // Finite-difference mode: when `frozen` is set, lobe selection, direction sampling and pdfs
    // use matsS/texelsS (the unperturbed parameters) while BSDF values and emission use
    // mats/texels. This is the function whose derivative the adjoint computes (sampling and pdfs
    // are detached in the reference, StandardBSDF.slang:155-178, PTUtils.slang:136-249).
    bool frozen = false;
    std::vector<float> matsS, texelsS;
    Camera cam;
    // Light collection.
    std::vector<EmissiveTri> lights;
    std::vector<uint32_t> triToLight, activeLights;
    // Oracle BVH2 (median split) over the triangles.
    struct Node { float lo[3], hi[3]; uint32_t left, right, first, count; };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    // Light BVH.
    struct LNode {
        V3f origin, extent;  // AABB center and half extent
        float flux;
        V3f coneDir;
        float cosCone;
        bool leaf;
        uint32_t right;      // internal: right child index (left = self + 1)
        uint32_t triOffset, triCount;
    };
    std::vector<LNode> lnodes;
    std::vector<uint32_t> lTriIdx;    // leaf triangle lists (indices into `lights`)
    std::vector<uint64_t> lBitmask;   // per light: traversal path bitmask
    // Analytic lights (LightData.slang:48-66 subset), 16 floats each after the setters' clamps:
    // [type (0 point, 1 directional), intensity rgb, posW, openingAngle, dirW, cosOpeningAngle,
    // penumbraAngle, 0, 0, 0].
    std::vector<float> aLights;
    // Environment map (EnvMap.slang): lat-long RGB texels, world-to-local rotation rows,
    // intensity * tint, flags (1 light, 2 background); the importance levels of EnvMapSampler,
    // r x r level at (r * r - 1) / 3 (empty: none, or zero luminance).
    std::vector<float> envTex;
    uint32_t envW = 0, envH = 0, envFlags = 0;
    V3f envInv[3], envScale;
    std::vector<float> envImp;
    uint32_t envImpDim = 0;

    void triVerts(uint32_t t, V3f v[3]) const {
        v[0] = pos[idx[3 * t]]; v[1] = pos[idx[3 * t + 1]]; v[2] = pos[idx[3 * t + 2]];
    }
    const float* mat(uint32_t m) const { return &mats[size_t(m) * M_STRIDE]; }
};

// ---------------------------------------------------------------- traversal
inline bool betterHit(float t, uint32_t id, const Hit& best) {
    return !best.valid() || t < best.t || (t == best.t && id < best.tri);
}
// Triangle test with the range rule the product also uses: tmin <= t <= tmax.
inline bool testTri(const Scene& s, uint32_t tri, V3f o, V3f d, float tmin, float tmax, Hit& best) {
    V3f v[3];
    s.triVerts(tri, v);
    V3f b; float t;
    if (!intersectRayTriangle(o, d, v, b, t)) return false;
    if (!(t >= tmin && t <= tmax)) return false;
    if (!betterHit(t, tri, best)) return false;
    best.tri = tri; best.t = t; best.u = b.y; best.v = b.z;
    return true;
}
inline Hit bruteClosest(const Scene& s, V3f o, V3f d, float tmin, float tmax) {
    Hit h;
    for (uint32_t i = 0; i < s.nT; i++) testTri(s, i, o, d, tmin, tmax, h);
    return h;
}

void buildOracleBVH(Scene& s);

// Double-precision slab test on a box grown by a relative margin; conservative for the
// float triangle test (a box is only culled when it is clearly beyond the current hit).
inline bool boxHit(const Scene::Node& n, const double o[3], const double inv[3], double tmax, double& tEntry) {
    double t0 = -1e300, t1 = 1e300;
    for (int a = 0; a < 3; a++) {
        double ext = std::max(std::fabs((double)n.lo[a]), std::fabs((double)n.hi[a])) * 1e-6 + 1e-30;
        double lo = (double)n.lo[a] - ext, hi = (double)n.hi[a] + ext;
        double ta = (lo - o[a]) * inv[a], tb = (hi - o[a]) * inv[a];
        if (std::isnan(ta) || std::isnan(tb)) { ta = -1e300; tb = 1e300; }
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta); t1 = std::min(t1, tb);
    }
    tEntry = t0;
    return t0 <= t1 * (1 + 1e-9) + 1e-30 && t1 >= -1e-30 && t0 <= tmax * (1 + 1e-9) + 1e-30;
}
// Closest hit; `anyHit` stops at the first accepted hit (visibility query).
inline Hit traceBVH(const Scene& s, V3f o, V3f d, float tmin, float tmax, bool anyHit, uint64_t* nodeCount = nullptr,
                    uint64_t* triCount = nullptr) {
    Hit best;
    if (s.nodes.empty()) return best;
    double od[3] = {o.x, o.y, o.z};
    double inv[3] = {1.0 / (double)d.x, 1.0 / (double)d.y, 1.0 / (double)d.z};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Scene::Node& n = s.nodes[stack[--sp]];
        double tE;
        double limit = best.valid() ? (double)best.t : (double)tmax;
        if (nodeCount) (*nodeCount)++;
        if (!boxHit(n, od, inv, limit, tE)) continue;
        if (n.count > 0) {
            for (uint32_t i = 0; i < n.count; i++) {
                if (triCount) (*triCount)++;
                float tm = best.valid() ? best.t : tmax;
                if (testTri(s, s.order[n.first + i], o, d, tmin, tm, best) && anyHit) return best;
            }
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return best;
}

// ---------------------------------------------------------------- shading data
// Scene.slang:440-476 + MaterialFactory.slang:78-104 with identity instance transforms
// (the scene format stores world-space geometry).
struct ShadingData {
    V3f posW, V, faceN;
    float uvx, uvy;
    Frame<float> frame;
    bool frontFacing;
    uint32_t mat;
};
inline ShadingData loadShadingData(const Scene& s, const Hit& h, V3f rayDir) {
    ShadingData sd;
    uint32_t i0 = s.idx[3 * h.tri], i1 = s.idx[3 * h.tri + 1], i2 = s.idx[3 * h.tri + 2];
    float b0 = 1.f - h.u - h.v, b1 = h.u, b2 = h.v;
    V3f p = s.pos[i0] * b0; p = p + s.pos[i1] * b1; p = p + s.pos[i2] * b2;
    V3f n = s.nrm[i0] * b0; n = n + s.nrm[i1] * b1; n = n + s.nrm[i2] * b2;
    float ux = s.uv[2 * i0] * b0; ux = ux + s.uv[2 * i1] * b1; ux = ux + s.uv[2 * i2] * b2;
    float uy = s.uv[2 * i0 + 1] * b0; uy = uy + s.uv[2 * i1 + 1] * b1; uy = uy + s.uv[2 * i2 + 1] * b2;
    V3f fn = normalize(cross(s.pos[i1] - s.pos[i0], s.pos[i2] - s.pos[i0]));
    n = normalize(n);
    sd.posW = p;
    sd.uvx = ux; sd.uvy = uy;
    sd.V = -rayDir;
    sd.frame = frameFromNormal(n);
    sd.faceN = fn;
    sd.frontFacing = dot(sd.V, sd.faceN) >= 0.f;
    sd.mat = s.triMat[h.tri];
    return sd;
}
inline V3f orientedFaceN(const ShadingData& sd) { return sd.frontFacing ? sd.faceN : -sd.faceN; }
inline V3f sdRayOrigin(const ShadingData& sd, bool viewside = true) {
    // ShadingData.slang:77-80
    return computeRayOrigin(sd.posW, (sd.frontFacing == viewside) ? sd.faceN : -sd.faceN);
}

// Bilinear texture fetch at mip 0 with wrap addressing (MaterialSystem.slang:308-320 with
// the default sampler of MaterialSystem.cpp:74-77). Also returns the 4 texel ids + weights.
struct TexFetch { uint32_t texel[4]; float w[4]; };
inline float wrapCoord(float x) { return x - std::floor(x); }
inline V3f sampleTexture(const Scene& s, int tex, float u, float v, TexFetch* tf = nullptr, bool sampling = false) {
    const int32_t* td = &s.texDesc[4 * tex];
    int W = td[1], H = td[2];
    float x = wrapCoord(u) * float(W) - 0.5f, y = wrapCoord(v) * float(H) - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float fx = x - fx0, fy = y - fy0;
    int x0 = int(fx0), y0 = int(fy0);
    int xa = ((x0 % W) + W) % W, xb = (((x0 + 1) % W) + W) % W;
    int ya = ((y0 % H) + H) % H, yb = (((y0 + 1) % H) + H) % H;
    const float* base = &((sampling && s.frozen) ? s.texelsS : s.texels)[td[0]];
    auto T = [&](int xx, int yy) { const float* q = base + 3 * (size_t(yy) * W + xx); return V3f(q[0], q[1], q[2]); };
    float wx0 = 1.f - fx, wy0 = 1.f - fy;
    V3f top = T(xa, ya) * wx0 + T(xb, ya) * fx;
    V3f bot = T(xa, yb) * wx0 + T(xb, yb) * fx;
    V3f c = top * wy0 + bot * fy;
    if (tf) {
        tf->texel[0] = uint32_t(ya * W + xa); tf->texel[1] = uint32_t(ya * W + xb);
        tf->texel[2] = uint32_t(yb * W + xa); tf->texel[3] = uint32_t(yb * W + xb);
        tf->w[0] = wx0 * wy0; tf->w[1] = fx * wy0; tf->w[2] = wx0 * fy; tf->w[3] = fx * fy;
    }
    return c;
}

// Material instance: StandardMaterial.slang:50-126 / PBRTDiffuseMaterial.slang.
struct MatInst {
    Frame<float> sf;
    BsdfData<float> d;   // evaluation parameters
    BsdfData<float> ds;  // sampling parameters (== d unless the scene is in frozen mode)
    V3f emission;
    StdBsdf<float> bsdf;  // initialised per wi
};
// MaterialSystem::evalEmissive (MaterialSystem.slang:201-213): emissive texture (bilinear, mip 0) or
// colour, times emissiveFactor.
inline V3f evalEmissive(const Scene& s, const float* m, float u, float v) {
    const int etex = int(m[M_EMIT_TEX]) - 1;
    V3f c = etex >= 0 ? sampleTexture(s, etex, u, v) : V3f(m[M_EMIT], m[M_EMIT + 1], m[M_EMIT + 2]);
    return c * m[M_EMIT_FACTOR];
}
// BasicMaterial::adjustDoubleSidedFlag (BasicMaterial.cpp:662-673): transmissive materials are
// double-sided.
inline bool matDoubleSided(const float* m) {
    return m[M_DOUBLE_SIDED] != 0.f ||
           (int(m[M_TYPE]) == MAT_STANDARD && (m[M_DIFF_TRANS] > 0.f || m[M_SPEC_TRANS] > 0.f));
}
// StandardMaterial.slang:56-63 (eta with the incident medium's IoR 1, transmission colour) and
// StandardBSDF.slang:131 (square root of the colour unless the surface is thin).
inline void setTransmission(BsdfData<float>& d, const float* m, const ShadingData& sd) {
    float ior = m[M_IOR];
    d.eta = sd.frontFacing ? (1.f / ior) : (ior / 1.f);
    d.diffTrans = m[M_DIFF_TRANS];
    d.specTrans = m[M_SPEC_TRANS];
    d.transmission = V3f(0.f);
    if (d.diffTrans > 0.f || d.specTrans > 0.f) {
        V3f t(m[M_TRANS], m[M_TRANS + 1], m[M_TRANS + 2]);
        d.transmission = m[M_THIN] != 0.f ? t : V3f(std::sqrt(t.x), std::sqrt(t.y), std::sqrt(t.z));
    }
}
inline BsdfData<float> bsdfDataFrom(const Scene& s, const float* m, const ShadingData& sd, bool sampling) {
    BsdfData<float> d;
    d.type = int(m[M_TYPE]);
    V3f base(m[M_BASE], m[M_BASE + 1], m[M_BASE + 2]);
    int tex = int(m[M_TEX]);
    if (tex >= 0) base = sampleTexture(s, tex, sd.uvx, sd.uvy, nullptr, sampling);
    if (d.type == MAT_PBRT_DIFFUSE) {
        d.diffuse = base; d.specular = V3f(0.f); d.roughness = 0.f; d.metallic = 0.f;
        return d;
    }
    setTransmission(d, m, sd);
    float ior = m[M_IOR], f = (ior - 1.f) / (ior + 1.f), F0 = f * f, metal = m[M_METAL];
    d.diffuse = base + (V3f(0.f) - base) * metal;
    d.specular = V3f(F0) + (base - V3f(F0)) * metal;
    d.roughness = m[M_ROUGH];
    d.metallic = metal;
    return d;
}
inline MatInst makeMatInst(const Scene& s, const ShadingData& sd) {
    MatInst mi;
    const float* m = s.mat(sd.mat);
    if (s.frozen) {
        mi.ds = bsdfDataFrom(s, &s.matsS[size_t(sd.mat) * M_STRIDE], sd, true);
        mi.d = bsdfDataFrom(s, m, sd, false);
        mi.sf = sd.frame;
        if (!sd.frontFacing && matDoubleSided(m)) mi.sf.N = -mi.sf.N;
        mi.emission = V3f(0.f);
        if (mi.d.type == MAT_STANDARD && sd.frontFacing) mi.emission = evalEmissive(s, m, sd.uvx, sd.uvy);
        return mi;
    }
    mi.d.type = int(m[M_TYPE]);
    V3f base(m[M_BASE], m[M_BASE + 1], m[M_BASE + 2]);
    int tex = int(m[M_TEX]);
    if (tex >= 0) base = sampleTexture(s, tex, sd.uvx, sd.uvy);
    mi.sf = sd.frame;
    // ShadingUtils.slang:107-117 (flipShadingNormal)
    if (!sd.frontFacing && matDoubleSided(m)) mi.sf.N = -mi.sf.N;
    if (mi.d.type == MAT_PBRT_DIFFUSE) {
        mi.d.diffuse = base;
        mi.d.specular = V3f(0.f);
        mi.d.roughness = 0.f; mi.d.metallic = 0.f;
        mi.emission = V3f(0.f);
        mi.ds = mi.d;
        return mi;
    }
    setTransmission(mi.d, m, sd);
    float ior = m[M_IOR];
    float f = (ior - 1.f) / (ior + 1.f);
    float F0 = f * f;
    float metal = m[M_METAL];
    // lerp(a, b, t) = a + (b - a) * t
    mi.d.diffuse = base + (V3f(0.f) - base) * metal;
    mi.d.specular = V3f(F0) + (base - V3f(F0)) * metal;
    mi.d.roughness = m[M_ROUGH];
    mi.d.metallic = metal;
    mi.emission = V3f(0.f);
    if (sd.frontFacing) mi.emission = evalEmissive(s, m, sd.uvx, sd.uvy);
    mi.ds = mi.d;
    return mi;
}
inline bool s_frozenGate(const MatInst& mi) {
    return mi.ds.type == MAT_STANDARD && (mi.ds.diffuse.x != mi.d.diffuse.x || mi.ds.diffuse.y != mi.d.diffuse.y ||
                                          mi.ds.diffuse.z != mi.d.diffuse.z || mi.ds.specular.x != mi.d.specular.x ||
                                          mi.ds.specular.y != mi.d.specular.y || mi.ds.specular.z != mi.d.specular.z ||
                                          mi.ds.roughness != mi.d.roughness || mi.ds.metallic != mi.d.metallic ||
                                          mi.ds.specTrans != mi.d.specTrans || mi.ds.diffTrans != mi.d.diffTrans);
}
// ShadingUtils.slang:155-181 (reflection) and :214-240 (reflection or transmission, fullSphere=false)
inline bool validHemisphere(const ShadingData& sd, V3f wiL, V3f woL, V3f wo) {
    bool isTransmission = wiL.z * woL.z < 0.f;
    if (std::min(wiL.z, std::fabs(woL.z)) < kMinCosTheta) return false;
    bool wiTop = sd.frontFacing;
    bool woTop = dot(wo, sd.faceN) >= 0.f;
    if (isTransmission) { if (wiTop == woTop) return false; }
    else { if (wiTop != woTop) return false; }
    return true;
}
inline bool validHemisphereReflection(const ShadingData& sd, V3f wiL, V3f woL, V3f wo) {
    if (std::min(wiL.z, woL.z) < kMinCosTheta) return false;
    bool wiTop = sd.frontFacing;
    bool woTop = dot(wo, sd.faceN) >= 0.f;
    return wiTop == woTop;
}
// StandardMaterialInstance.slang:63-74 / PBRTDiffuseMaterialInstance.slang:54-64
inline V3f miEval(const ShadingData& sd, MatInst& mi, V3f wo) {
    V3f wiL = mi.sf.toLocal(sd.V), woL = mi.sf.toLocal(wo);
    if (mi.d.type == MAT_PBRT_DIFFUSE) {
        if (!validHemisphereReflection(sd, wiL, woL, wo)) return V3f(0.f);
        return mi.d.diffuse * std::max(woL.z, 0.0f) * Consts<float>::inv_pi;
    }
    if (!validHemisphere(sd, wiL, woL, wo)) return V3f(0.f);
    mi.bsdf.init(wiL, mi.d);
    if (s_frozenGate(mi)) {
        StdBsdf<float> g;
        g.init(wiL, mi.ds);
        mi.bsdf.pDiff = g.pDiff;
        mi.bsdf.pDiffT = g.pDiffT;
        mi.bsdf.pSpec = g.pSpec;
        mi.bsdf.pSpecT = g.pSpecT;
    }
    return mi.bsdf.eval(wiL, woL);
}
inline float miEvalPdf(const ShadingData& sd, MatInst& mi, V3f wo) {
    V3f wiL = mi.sf.toLocal(sd.V), woL = mi.sf.toLocal(wo);
    if (mi.d.type == MAT_PBRT_DIFFUSE) return std::max(woL.z, 0.0f) * Consts<float>::inv_pi;
    if (!validHemisphere(sd, wiL, woL, wo)) return 0.f;
    mi.bsdf.init(wiL, mi.ds);
    return mi.bsdf.evalPdf(wiL, woL);
}
// transmission: the sampled lobe is a transmission lobe (LobeType::Transmission), i.e. the local
// direction lies below the shading hemisphere.
struct BsdfSample { V3f wo, weight; float pdf; bool delta, transmission = false; };
// StandardMaterialInstance.slang:89-110 (+ sampleReference :202-231) / PBRTDiffuseMaterialInstance.slang:66-78
inline bool miSample(const ShadingData& sd, MatInst& mi, SampleGen& sg, BsdfSample& bs, bool importance = true) {
    V3f wiL = mi.sf.toLocal(sd.V);
    V3f woL;
    bs.delta = false;
    if (mi.d.type == MAT_PBRT_DIFFUSE) {
        float ux = sg.next1D(), uy = sg.next1D();
        woL = sample_cosine_hemisphere_concentric(ux, uy, bs.pdf);
        bs.weight = mi.ds.diffuse;
        bs.wo = mi.sf.fromLocal(woL);
        if (!validHemisphereReflection(sd, wiL, woL, bs.wo) || bs.pdf == 0.f) return false;
        return true;
    }
    if (!importance) {
        // sampleReference (StandardMaterialInstance.slang:202-231): mirrored below the surface with
        // probability 1/2 when the lobe set holds a transmission lobe (StandardBSDF.slang:205-225).
        const BsdfData<float>& q = mi.ds;
        bool anyDiffuse = q.diffuse.x > 0.f || q.diffuse.y > 0.f || q.diffuse.z > 0.f;
        bool transmissive = q.specTrans > 0.f || (anyDiffuse && q.specTrans < 1.f && q.diffTrans > 0.f);
        float ux = sg.next1D(), uy = sg.next1D();
        woL = sample_cosine_hemisphere_concentric(ux, uy, bs.pdf);
        if (transmissive) {
            if (sg.next1D() < 0.5f) woL.z = -woL.z;
            bs.pdf *= 0.5f;
            if (std::min(std::fabs(wiL.z), std::fabs(woL.z)) < kMinCosTheta || bs.pdf == 0.f) return false;
        } else if (std::min(wiL.z, woL.z) < kMinCosTheta || bs.pdf == 0.f) {
            return false;
        }
        mi.bsdf.init(wiL, mi.ds);
        bs.weight = mi.bsdf.eval(wiL, woL) / bs.pdf;
    } else {
        mi.bsdf.init(wiL, mi.ds);
        if (!mi.bsdf.sample(wiL, woL, bs.pdf, bs.weight, bs.delta, sg)) return false;
    }
    bs.wo = mi.sf.fromLocal(woL);
    bs.transmission = woL.z < 0.f;
    if (!validHemisphere(sd, wiL, woL, bs.wo)) return false;
    return true;
}

// ---------------------------------------------------------------- light collection
void buildLightCollection(Scene& s);
void updateLightPositions(Scene& s);
void refitLightBVH(Scene& s);
void buildLightBVH(Scene& s);
// Precision of the packed light-BVH node fields (fo_impl.cpp): half-float extents, the quantised
// cone cosine, the octahedral 2 x 16-bit cone axis.
float halfPrecision(float v);
float conePrecision(float c);
V3f octNormalPrecision(V3f n);

struct LightSample { V3f posW, normalW, Le, dir; float pdf, distance; uint32_t light; };
// EmissiveLightSamplerHelpers.slang:54-101
inline bool sampleTriangle(const Scene& s, V3f posW, uint32_t li, float ux, float uy, LightSample& ls) {
    const EmissiveTri& tri = s.lights[li];
    ls.light = li;
    V3f b = sample_triangle(ux, uy);
    ls.posW = tri.p[0] * b.x + tri.p[1] * b.y + tri.p[2] * b.z;
    V3f toLight = ls.posW - posW;
    const float distSqr = std::max(std::numeric_limits<float>::min(), dot(toLight, toLight));
    ls.distance = std::sqrt(distSqr);
    ls.dir = toLight / ls.distance;
    ls.normalW = tri.normal;
    float cosTheta = dot(ls.normalW, -ls.dir);
    if (cosTheta <= 0.f) return false;
    // EmissiveLightSamplerHelpers.slang:79: evalEmissive of the triangle's current material.
    const float u = tri.uv[0][0] * b.x + tri.uv[1][0] * b.y + tri.uv[2][0] * b.z;
    const float v = tri.uv[0][1] * b.x + tri.uv[1][1] * b.y + tri.uv[2][1] * b.z;
    ls.Le = evalEmissive(s, s.mat(s.triMat[tri.tri]), u, v);
    float denom = std::max(std::numeric_limits<float>::min(), cosTheta * tri.area);
    ls.pdf = distSqr / denom;
    return true;
}
// EmissiveLightSamplerHelpers.slang:106-124
inline float evalTrianglePdf(const Scene& s, V3f posW, uint32_t li, V3f hitPos, V3f hitNormal) {
    V3f toLight = hitPos - posW;
    const float distSqr = dot(toLight, toLight);
    if (distSqr <= std::numeric_limits<float>::min()) return 0.f;
    V3f L = toLight / std::sqrt(distSqr);
    float cosTheta = dot(hitNormal, -L);
    if (cosTheta <= 0.f) return 0.f;
    float denom = std::max(std::numeric_limits<float>::min(), cosTheta * s.lights[li].area);
    return distSqr / denom;
}
// LightBVHSampler.slang:219-273 (Sphere bound, bounding + lighting cones, node flux).
float lbvhImportance(const Scene& s, V3f posW, V3f n, uint32_t node);
bool lbvhSample(const Scene& s, V3f posW, V3f n, float u, float& pdf, uint32_t& light);
float lbvhSelectionPdf(const Scene& s, V3f posW, V3f n, uint32_t light);

enum LightSamplerKind { LS_UNIFORM = 0, LS_LIGHTBVH = 1 };
// EmissiveUniformSampler.slang:55-74 and LightBVHSampler.slang:75-94.
inline bool sampleLight(const Scene& s, int kind, V3f posW, V3f n, SampleGen& sg, LightSample& ls) {
    if (s.activeLights.empty()) return false;
    float uLight = sg.next1D();
    uint32_t li;
    float selPdf;
    if (kind == LS_UNIFORM) {
        uint32_t cnt = uint32_t(s.activeLights.size());
        uint32_t idx = std::min(uint32_t(uLight * float(cnt)), cnt - 1);
        li = s.activeLights[idx];
        selPdf = 1.f / float(cnt);
    } else {
        if (!lbvhSample(s, posW, n, uLight, selPdf, li)) return false;
    }
    float ux = sg.next1D(), uy = sg.next1D();
    if (!sampleTriangle(s, posW, li, ux, uy, ls)) return false;
    ls.pdf *= selPdf;
    return true;
}
inline float lightPdf(const Scene& s, int kind, V3f posW, V3f n, uint32_t li, V3f hitPos, V3f hitNormal) {
    if (li == kInvalid) return 0.f;
    float sel;
    if (kind == LS_UNIFORM) sel = s.activeLights.empty() ? 0.f : 1.f / float(s.activeLights.size());
    else sel = lbvhSelectionPdf(s, posW, n, li);
    return sel * evalTrianglePdf(s, posW, li, hitPos, hitNormal);
}
// PTUtils.slang:87-93 (power heuristic)
inline float evalMIS(float n0, float p0, float n1, float p1) {
    float q0 = (n0 * p0) * (n0 * p0);
    float q1 = (n1 * p1) * (n1 * p1);
    return q0 / (q0 + q1);
}

// ---------------------------------------------------------------- environment map
// Bilinear lat-long lookup at mip 0, wrap in u, clamp in v (EnvMap.cpp:131-135).
inline V3f envLookup(const Scene& s, float u, float v) {
    const int W = int(s.envW), H = int(s.envH);
    const float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
    const float x0f = std::floor(x), y0f = std::floor(y);
    const float fx = x - x0f, fy = y - y0f;
    const int x0 = int(x0f), y0 = int(y0f);
    auto wrap = [](int a, int m) { int r = a % m; return r < 0 ? r + m : r; };
    auto clampi = [](int a, int m) { return a < 0 ? 0 : (a >= m ? m - 1 : a); };
    const int xa = wrap(x0, W), xb = wrap(x0 + 1, W), ya = clampi(y0, H), yb = clampi(y0 + 1, H);
    auto T = [&](int xx, int yy) { const float* q = &s.envTex[3 * (size_t(yy) * W + xx)]; return V3f(q[0], q[1], q[2]); };
    const V3f top = T(xa, ya) * (1.f - fx) + T(xb, ya) * fx;
    const V3f bot = T(xa, yb) * (1.f - fx) + T(xb, yb) * fx;
    return top * (1.f - fy) + bot * fy;
}
// world_to_latlong_map (MathHelpers.slang:111-118).
inline void latlongMap(V3f dir, float& u, float& v) {
    const V3f p = normalize(dir);
    u = std::atan2(p.x, -p.z) * (0.5f / Consts<float>::pi) + 0.5f;
    v = std::acos(std::min(std::max(p.y, -1.f), 1.f)) / Consts<float>::pi;
}
inline V3f envToLocal(const Scene& s, V3f d) { return V3f(dot(s.envInv[0], d), dot(s.envInv[1], d), dot(s.envInv[2], d)); }
inline V3f envToWorld(const Scene& s, V3f l) { return s.envInv[0] * l.x + s.envInv[1] * l.y + s.envInv[2] * l.z; }
// EnvMap::eval (EnvMap.slang:58-68).
inline V3f envEval(const Scene& s, V3f dir) {
    float u, v;
    latlongMap(envToLocal(s, dir), u, v);
    const V3f c = envLookup(s, u, v);
    return V3f(c.x * s.envScale.x, c.y * s.envScale.y, c.z * s.envScale.z);
}
inline float signOf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
// oct_to_ndir_equal_area_unorm / ndir_to_oct_equal_area_unorm (MathHelpers.slang:211-254).
inline V3f octToDirEqualArea(float px, float py) {
    px = px * 2.f - 1.f;
    py = py * 2.f - 1.f;
    const float d = 1.f - (std::fabs(px) + std::fabs(py));
    const float r = 1.f - std::fabs(d);
    const float phi = r > 0.f ? ((std::fabs(py) - std::fabs(px)) / r + 1.f) * (Consts<float>::pi / 4.f) : 0.f;
    const float f = r * std::sqrt(2.f - r * r);
    return V3f(f * signOf(px) * std::cos(phi), f * signOf(py) * std::sin(phi), signOf(d) * (1.f - r * r));
}
inline void dirToOctEqualArea(V3f n, float& ux, float& uy) {
    const float r = std::sqrt(1.f - std::fabs(n.z));
    const float phi = std::atan2(std::fabs(n.y), std::fabs(n.x));
    float py = r * phi * (2.f / Consts<float>::pi);
    float px = r - py;
    if (n.z < 0.f) { const float t = px; px = 1.f - py; py = 1.f - t; }
    px *= signOf(n.x);
    py *= signOf(n.y);
    ux = px * 0.5f + 0.5f;
    uy = py * 0.5f + 0.5f;
}
// EnvMapSamplerSetup.cs.slang:44-68 for one texel of the dim x dim level, ns x ns samples.
inline float envImportanceTexel(const Scene& s, uint32_t x, uint32_t y, uint32_t dim, uint32_t ns) {
    float L = 0.f;
    for (uint32_t sy = 0; sy < ns; sy++)
        for (uint32_t sx = 0; sx < ns; sx++) {
            const float px = (float(x * ns + sx) + 0.5f) / float(dim * ns), py = (float(y * ns + sy) + 0.5f) / float(dim * ns);
            float u, v;
            latlongMap(octToDirEqualArea(px, py), u, v);
            L += luminance(envLookup(s, u, v));
        }
    return L / float(ns * ns);
}
// EnvMapSampler::sample (EnvMapSampler.slang:73-145).
inline bool envSample(const Scene& s, float ux, float uy, V3f& dir, float& pdf, V3f& Le) {
    const uint32_t dim = s.envImpDim;
    uint32_t px = 0, py = 0;
    for (uint32_t r = 2; r <= dim; r *= 2) {
        px *= 2; py *= 2;
        const float* L = &s.envImp[(r * r - 1) / 3];
        const float w0 = L[py * r + px], w1 = L[py * r + px + 1], w2 = L[(py + 1) * r + px], w3 = L[(py + 1) * r + px + 1];
        const float q0 = w0 + w2, q1 = w1 + w3;
        const float d = q0 / (q0 + q1);
        uint32_t ox, oy;
        if (ux < d) { ox = 0; ux = ux / d; } else { ox = 1; ux = (ux - d) / (1.f - d); }
        const float e = ox == 0 ? w0 / q0 : w1 / q1;
        if (uy < e) { oy = 0; uy = uy / e; } else { oy = 1; uy = (uy - e) / (1.f - e); }
        px += ox; py += oy;
    }
    const V3f l = octToDirEqualArea((float(px) + ux) / float(dim), (float(py) + uy) / float(dim));
    pdf = s.envImp[(dim * dim - 1) / 3 + py * dim + px] / s.envImp[0] * (0.25f / Consts<float>::pi);
    dir = envToWorld(s, l);
    Le = envEval(s, dir);
    return true;
}
// EnvMapSampler::evalPdf (EnvMapSampler.slang:154-163): point sampling with clamping.
inline float envPdf(const Scene& s, V3f dir) {
    float ux, uy;
    dirToOctEqualArea(envToLocal(s, dir), ux, uy);
    const int dim = int(s.envImpDim);
    const int x = std::min(std::max(int(std::floor(ux * float(dim))), 0), dim - 1);
    const int y = std::min(std::max(int(std::floor(uy * float(dim))), 0), dim - 1);
    return s.envImp[(size_t(dim) * dim - 1) / 3 + size_t(y) * dim + x] / s.envImp[0] * (0.25f / Consts<float>::pi);
}

// ---------------------------------------------------------------- integrators
struct TraceStats { uint64_t closest = 0, visibility = 0, nodes = 0, tris = 0; };
struct RayRecord { V3f o, d; float tmin, tmax; uint32_t tri; float t, u, v; uint32_t kind; };

struct MptParams {
    uint32_t maxBounces = 3;
    bool computeDirect = true;
    bool useImportanceSampling = true;
    bool useEmissiveNEE = false;  // extension: emissive next-event estimation with MIS
    bool useEnvNEE = false;       // extension: env map next-event estimation with MIS
    int lightSampler = LS_LIGHTBVH;
    bool brute = false;           // brute-force traversal (small scenes)
};

struct Tracer {
    const Scene& s;
    bool brute;
    TraceStats* stats;
    std::vector<RayRecord>* rec;
    Hit closest(V3f o, V3f d, float tmin, float tmax) {
        if (stats) stats->closest++;
        Hit h = brute ? bruteClosest(s, o, d, tmin, tmax)
                      : traceBVH(s, o, d, tmin, tmax, false, stats ? &stats->nodes : nullptr, stats ? &stats->tris : nullptr);
        if (rec) rec->push_back({o, d, tmin, tmax, h.tri, h.t, h.u, h.v, 0});
        return h;
    }
    bool visible(V3f o, V3f d, float tmax) {
        if (stats) stats->visibility++;
        Hit h;
        if (brute) {
            for (uint32_t i = 0; i < s.nT && !h.valid(); i++) testTri(s, i, o, d, 0.f, tmax, h);
        } else {
            h = traceBVH(s, o, d, 0.f, tmax, true, stats ? &stats->nodes : nullptr, stats ? &stats->tris : nullptr);
        }
        if (rec) rec->push_back({o, d, 0.f, tmax, h.tri, h.t, h.u, h.v, 1});
        return !h.valid();
    }
};

// Primary ray of a pinhole camera, pixel centre (Camera.slang:46-77).
inline void primaryRay(const Camera& c, uint32_t px, uint32_t py, uint32_t W, uint32_t H, V3f& o, V3f& d, float& tmin, float& tmax) {
    float p_x = (float(px) + 0.5f) / float(W), p_y = (float(py) + 0.5f) / float(H);
    float ndc_x = 2.f * p_x + -1.f, ndc_y = -2.f * p_y + 1.f;
    V3f dir = c.U * ndc_x + c.V * ndc_y + c.W;
    d = normalize(dir);
    o = c.pos;
    float invCos = 1.f / dot(normalize(c.W), d);
    tmin = c.nearZ * invCos;
    tmax = c.farZ * invCos;
}

// NEE extension (shared by both integrator modes of this oracle): one emissive light sample
// with a shadow ray; MIS against the BSDF pdf (PTUtils.slang:177-249 without reparameterization).
// `sel` is the probability of having chosen emissive lights among the light types.
inline V3f neeEmissive(const Scene& s, const MptParams& P, const ShadingData& sd, MatInst& mi, SampleGen& sg, Tracer& tr,
                       float sel = 1.f) {
    LightSample ls;
    V3f nrm = mi.sf.N;
    if (!sampleLight(s, P.lightSampler, sd.posW, nrm, sg, ls)) return V3f(0.f);
    if (!(ls.pdf > 0.f)) return V3f(0.f);
    V3f Li = ls.Le / ls.pdf;
    if (!anyPositive(Li)) return V3f(0.f);
    ls.pdf *= sel;  // generateLightSample (PathTracer.slang:729-732)
    Li = Li / sel;
    V3f origin = computeRayOrigin(sd.posW, dot(sd.faceN, ls.dir) >= 0.f ? sd.faceN : -sd.faceN);
    V3f lightPos = computeRayOrigin(ls.posW, ls.normalW);
    V3f toLight = lightPos - origin;
    float dist = length(toLight);
    V3f dir = normalize(toLight);
    V3f f = miEval(sd, mi, dir);
    if (!anyPositive(f)) return V3f(0.f);
    float scatterPdf = miEvalPdf(sd, mi, dir);
    float mis = evalMIS(1.f, ls.pdf, 1.f, scatterPdf);
    V3f contrib = f * Li * mis;
    if (!tr.visible(origin, dir, dist)) return V3f(0.f);
    return contrib;
}

// MinimalPathTracer.rt.slang:253-286 (evalDirectAnalytic) with samplePointLight /
// sampleDirectionalLight (LightHelpers.slang:190-258).
inline V3f evalDirectAnalytic(const Scene& s, const ShadingData& sd, MatInst& mi, SampleGen& sg, Tracer& tr) {
    const uint32_t n = uint32_t(s.aLights.size() / 16);
    if (n == 0) return V3f(0.f);
    const uint32_t li = std::min(uint32_t(sg.next1D() * float(n)), n - 1);
    const float invPdf = float(n);
    const float* L = &s.aLights[16 * size_t(li)];
    const V3f intensity(L[1], L[2], L[3]), dirW(L[8], L[9], L[10]);
    V3f dir, Li;
    float distance;
    if (L[0] == 1.f) {
        dir = -dirW;
        distance = std::numeric_limits<float>::max();
        Li = intensity;
    } else {
        V3f toLight = V3f(L[4], L[5], L[6]) - sd.posW;
        float distSqr = std::max(dot(toLight, toLight), 1e-9f);  // kMinLightDistSqr
        distance = std::sqrt(distSqr);
        dir = toLight / distance;
        float cosTheta = -dot(dir, dirW);
        float falloff = 1.f;
        if (cosTheta < L[11]) {
            falloff = 0.f;
        } else if (L[12] > 0.f) {
            float delta = L[7] - std::acos(cosTheta);
            float t = std::min(std::max((delta - 0.f) / (L[12] - 0.f), 0.f), 1.f);  // smoothstep(0, penumbra, delta)
            falloff = t * t * (3.f - 2.f * t);
        }
        Li = intensity * falloff / distSqr;
    }
    // getLobeTypes: a reflection lobe always exists; transmission as StandardBSDF.slang:205-225.
    const BsdfData<float>& q = mi.d;
    const bool anyDiffuse = q.diffuse.x > 0.f || q.diffuse.y > 0.f || q.diffuse.z > 0.f;
    const bool hasTransmission = q.type == MAT_STANDARD && (q.specTrans > 0.f || (anyDiffuse && q.specTrans < 1.f && q.diffTrans > 0.f));
    const float NdotL = dot(orientedFaceN(sd), dir);
    if (NdotL <= kMinCosTheta && !hasTransmission) return V3f(0.f);
    const V3f origin = computeRayOrigin(sd.posW, dot(sd.faceN, dir) >= 0.f ? sd.faceN : -sd.faceN);
    if (!tr.visible(origin, dir, distance)) return V3f(0.f);
    return miEval(sd, mi, dir) * Li * invPdf;
}

// Env map NEE (PathTracer.slang generateEnvMapSample, NEE with MIS): one EnvMapSampler sample
// and a visibility ray to infinity.
inline V3f neeEnvMap(const Scene& s, const ShadingData& sd, MatInst& mi, SampleGen& sg, Tracer& tr, float sel) {
    const float ux = sg.next1D(), uy = sg.next1D();
    V3f dir, Le;
    float pdf;
    if (!envSample(s, ux, uy, dir, pdf, Le)) return V3f(0.f);
    V3f Li = pdf > 0.f ? Le / pdf : V3f(0.f);
    if (!anyPositive(Li)) return V3f(0.f);
    pdf *= sel;
    Li = Li / sel;
    const V3f origin = computeRayOrigin(sd.posW, dot(sd.faceN, dir) >= 0.f ? sd.faceN : -sd.faceN);
    const V3f f = miEval(sd, mi, dir);
    if (!anyPositive(f)) return V3f(0.f);
    const V3f contrib = f * Li * evalMIS(1.f, pdf, 1.f, miEvalPdf(sd, mi, dir));
    if (!tr.visible(origin, dir, std::numeric_limits<float>::max())) return V3f(0.f);
    return contrib;
}

// Light-type selection probabilities (PathTracer.slang:643-663): equal among env map (when
// useEnvNEE and a sampleable env light exist), emissive (useEmissiveNEE), analytic (any).
struct LightTypeProbs { float env = 0.f, emissive = 0.f, analytic = 0.f; uint32_t count = 0; };
inline LightTypeProbs lightTypeProbs(const Scene& s, const MptParams& P) {
    const bool env = P.useEnvNEE && s.envImpDim > 0 && (s.envFlags & 1u);
    const bool em = P.useEmissiveNEE, an = !s.aLights.empty();
    LightTypeProbs t;
    t.count = uint32_t(env) + uint32_t(em) + uint32_t(an);
    const float inv = t.count ? 1.f / float(t.count) : 0.f;
    t.env = env ? inv : 0.f; t.emissive = em ? inv : 0.f; t.analytic = an ? inv : 0.f;
    return t;
}

// Direct light at a vertex: selectLightType (PathTracer.slang:671-693) and the chosen type's
// sample with one visibility ray. The selection draws a number only when more than one type is
// enabled; analytic lights alone are evalDirectAnalytic.
inline V3f directLight(const Scene& s, const MptParams& P, const ShadingData& sd, MatInst& mi, SampleGen& sg, Tracer& tr) {
    const LightTypeProbs t = lightTypeProbs(s, P);
    if (t.count == 0) return V3f(0.f);
    int type;
    float sel = 1.f;
    if (t.count == 1) {
        type = t.env > 0.f ? 0 : (t.emissive > 0.f ? 1 : 2);
    } else {
        float u = sg.next1D();
        if (u < t.env) { type = 0; sel = t.env; }
        else if ((u -= t.env) < t.emissive) { type = 1; sel = t.emissive; }
        else if ((u -= t.emissive) < t.analytic) { type = 2; sel = t.analytic; }
        else return V3f(0.f);
    }
    if (type == 0) return neeEnvMap(s, sd, mi, sg, tr, sel);
    if (type == 1) return neeEmissive(s, P, sd, mi, sg, tr, sel);
    return evalDirectAnalytic(s, sd, mi, sg, tr) / sel;
}

// MinimalPathTracer.rt.slang:392-468 (one path for one pixel and sample number).
inline V3f mptTracePath(const Scene& s, const MptParams& P, uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                        uint32_t frameCount, Tracer& tr) {
    V3f o, d; float tmin, tmax;
    primaryRay(s.cam, px, py, W, H, o, d, tmin, tmax);
    Hit hit = tr.closest(o, d, tmin, tmax);
    if (!hit.valid()) return !s.envTex.empty() && (s.envFlags & 2u) ? envEval(s, d) : V3f(0.f);  // :462-465
    ShadingData sd = loadShadingData(s, hit, d);
    MatInst mi = makeMatInst(s, sd);
    SampleGen sg = makeSampleGen(px, py, frameCount);
    V3f rayOrigin = sdRayOrigin(sd);
    V3f outColor(0.f);
    if (P.computeDirect) {
        outColor = outColor + mi.emission;
        outColor = outColor + directLight(s, P, sd, mi, sg, tr);
    }
    // ScatterRayData
    V3f radiance(0.f), thp(1.f), rOrigin, rDir;
    uint32_t pathLength = 0;
    bool terminated = false;
    float prevPdf = 0.f; bool prevDelta = false; V3f prevN; V3f prevPos;
    auto scatter = [&](const ShadingData& sdd, MatInst& mii, V3f org) -> bool {
        BsdfSample bs;
        if (miSample(sdd, mii, sg, bs, P.useImportanceSampling)) {
            // MinimalPathTracer.rt.slang:303-307: transmitted rays leave from the far side.
            rOrigin = bs.transmission ? sdRayOrigin(sdd, false) : org;
            rDir = bs.wo;
            thp = thp * bs.weight;
            prevPdf = bs.pdf; prevDelta = bs.delta; prevN = mii.sf.N; prevPos = rOrigin;
            return anyPositive(thp);
        }
        return false;
    };
    if (!scatter(sd, mi, rayOrigin)) terminated = true;
    for (uint32_t depth = 0; depth <= P.maxBounces && !terminated; depth++) {
        Hit h = tr.closest(rOrigin, rDir, 0.f, std::numeric_limits<float>::max());
        if (!h.valid()) {
            // scatterMiss (MinimalPathTracer.rt.slang:475-486), with MIS against env map sampling
            // at the previous vertex (PathTracer.slang:1096-1108).
            if (!s.envTex.empty() && (s.envFlags & 1u) && (P.computeDirect || pathLength > 0)) {
                V3f Le = envEval(s, rDir);
                const float pEnv = lightTypeProbs(s, P).env;
                if (pEnv > 0.f && !prevDelta) Le = Le * evalMIS(1.f, prevPdf, 1.f, envPdf(s, rDir) * pEnv);
                radiance = radiance + thp * Le;
            }
            terminated = true;
            break;
        }
        ShadingData sdh = loadShadingData(s, h, rDir);
        MatInst mih = makeMatInst(s, sdh);
        if (P.computeDirect || pathLength > 0) {
            V3f e = mih.emission;
            if (P.useEmissiveNEE && anyPositive(e) && !prevDelta) {
                uint32_t li = s.triToLight[h.tri];
                float lp = lightPdf(s, P.lightSampler, prevPos, prevN, li, sdh.posW, orientedFaceN(sdh));
                e = e * evalMIS(1.f, prevPdf, 1.f, lp * lightTypeProbs(s, P).emissive);
            }
            radiance = radiance + thp * e;
        }
        if (pathLength >= P.maxBounces) { terminated = true; break; }
        V3f org = sdRayOrigin(sdh);
        if (lightTypeProbs(s, P).count) radiance = radiance + thp * directLight(s, P, sdh, mih, sg, tr);
        if (!scatter(sdh, mih, org)) { terminated = true; break; }
        pathLength++;
    }
    return outColor + radiance;
}

// WARDiffPathTracer primal (useWAR = false): WARDiffPathTracer.rt.slang:75-143 and
// PTUtils.slang:97-341, uniform emissive sampler, NEE + MIS, BSDF sampling.
struct WarParams {
    uint32_t maxBounces = 0;
    bool useBSDFSampling = true, useNEE = true, useMIS = true, useAntithetic = true;
    bool brute = false;
    int lightSampler = LS_UNIFORM;
};
void computeIntersectionAD(const Scene& s, const Hit& h, V3f o, V3f d, V3f& posW, V3f& bary, V3f& faceN, V3f& normalW);
V3f warTracePaths(const Scene& s, const WarParams& P, uint32_t px, uint32_t py, uint32_t W, uint32_t H, uint32_t frameSeed, Tracer& tr);

}  // namespace fo
